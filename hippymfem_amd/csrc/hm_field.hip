// The seeded inputs of the serving form, drawn on the device from a seed held in device
// memory, so that a captured HIP graph of a whole request reads a new seed at every replay.
//
// Normal deviates: SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) hashes (key, index) into uniform words, and the Box-Muller
// transform (Box and Muller, 1958) turns two of them into one N(0, 1) value.  Every value
// depends on its index only, so the draw is the same in any launch shape.
//
// The parameter field is the benchmarks' one: N(0, 1) noise on the (n+1)^3 vertex lattice,
// three passes of the 7-point average over the neighbours that exist, then shifted and scaled
// to mean 0 and population standard deviation 1 (divisor n, as numpy's std).  The mean and
// deviation are sums in a fixed order (one block, a fixed tree), so the field is the same bits
// at every draw.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hm_common.h"
#include "hm_device.h"

namespace hm {

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in (0, 1): the top 53 bits, offset by half a step
__device__ inline double unit(uint64_t h) { return ((double)(h >> 11) + 0.5) * 0x1.0p-53; }

__device__ inline double normal_at(uint64_t key, int64_t i) {
  const double u1 = unit(splitmix64(key ^ (2 * (uint64_t)i)));
  const double u2 = unit(splitmix64(key ^ (2 * (uint64_t)i + 1)));
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ inline uint64_t key_of(int64_t seed, int64_t salt) {
  return splitmix64(splitmix64((uint64_t)seed) ^ (uint64_t)salt);
}

__global__ void k_normal(const int64_t* __restrict__ seed, int64_t salt, int64_t n, double* __restrict__ out) {
  const uint64_t key = key_of(*seed, salt);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = normal_at(key, i);
}

// one pass of the 7-point average over the neighbours that exist (the lattice is n1^3, x fastest)
__global__ void k_smooth(const double* __restrict__ f, int n1, double* __restrict__ g) {
  const int64_t n = (int64_t)n1 * n1 * n1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % n1), y = (int)((i / n1) % n1), z = (int)(i / ((int64_t)n1 * n1));
    const int64_t sy = n1, sz = (int64_t)n1 * n1;
    double acc = f[i];
    int cnt = 1;
    if (x > 0) acc += f[i - 1], ++cnt;
    if (x < n1 - 1) acc += f[i + 1], ++cnt;
    if (y > 0) acc += f[i - sy], ++cnt;
    if (y < n1 - 1) acc += f[i + sy], ++cnt;
    if (z > 0) acc += f[i - sz], ++cnt;
    if (z < n1 - 1) acc += f[i + sz], ++cnt;
    g[i] = acc / cnt;
  }
}

constexpr int FS_THREADS = 1024;

// Sum of v[i] - shift over i < n by one block: each thread its strided items in order, then a
// fixed tree in LDS.  Every thread gets the sum.
__device__ double block_sum(const double* __restrict__ v, int64_t n, double shift, bool square, double* red) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += FS_THREADS) {
    const double d = v[i] - shift;
    s += square ? d * d : d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = FS_THREADS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const double total = red[0];
  __syncthreads();
  return total;
}

// stats[0] = mean, stats[1] = population standard deviation (divisor n, numpy's np.std)
__global__ void __launch_bounds__(FS_THREADS) k_field_stats(const double* __restrict__ f, int64_t n, double* stats) {
  __shared__ double red[FS_THREADS];
  const double mean = block_sum(f, n, 0.0, false, red) / (double)n;
  const double ss = block_sum(f, n, mean, true, red);
  if (threadIdx.x == 0) {
    stats[0] = mean;
    stats[1] = n > 0 && ss > 0.0 ? sqrt(ss / (double)n) : 1.0;
  }
}

// out[to[i]] (or out[i]) = (f[i] - mean) / std
__global__ void k_field_finish(const double* __restrict__ f, int64_t n, const double* __restrict__ stats,
                               const int64_t* __restrict__ to, double* __restrict__ out) {
  const double mean = stats[0], sd = stats[1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[to ? to[i] : i] = (f[i] - mean) / sd;
}

static unsigned grid_for(int64_t n) {
  const int64_t g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g < 4096 ? g : 4096));
}

// the draw into a workspace of 2 n1^3 + 2 doubles: two lattice buffers and the statistics
static int field_into(const int64_t* seed, int n1, int passes, double* work, const int64_t* to, double* out,
                      hipStream_t st, const char* what) {
  const int64_t n = (int64_t)n1 * n1 * n1;
  double* a = work;
  double* b = work + n;
  double* stats = work + 2 * n;
  hipLaunchKernelGGL(k_normal, dim3(grid_for(n)), dim3(256), 0, st, seed, (int64_t)0, n, a);
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_smooth, dim3(grid_for(n)), dim3(256), 0, st, (const double*)a, n1, b);
    double* t = a;
    a = b;
    b = t;
  }
  hipLaunchKernelGGL(k_field_stats, dim3(1), dim3(FS_THREADS), 0, st, (const double*)a, n, stats);
  hipLaunchKernelGGL(k_field_finish, dim3(grid_for(n)), dim3(256), 0, st, (const double*)a, n,
                     (const double*)stats, to, out);
  return hip_check(what);
}

// The workspace of hm_serving_field_noalloc: one array a device, held until the process ends.
// Host code reaches it only through its device address (hipGetSymbolAddress, once a device).
constexpr int FS_NOALLOC_MAX_N1 = HM_SERVING_FIELD_NOALLOC_MAX_N1;
constexpr int64_t FS_NOALLOC_WORDS =
    2 * (int64_t)FS_NOALLOC_MAX_N1 * FS_NOALLOC_MAX_N1 * FS_NOALLOC_MAX_N1 + 2;
static __device__ __align__(256) double scratch[FS_NOALLOC_WORDS];

constexpr int FS_MAX_DEVICES = 64;
static std::mutex scratch_mutex;
static double* scratch_ptr[FS_MAX_DEVICES];

// the device address of the current device's workspace (looked up on its first call)
static int scratch_of_device(double** out) {
  int dev = 0;
  const hipError_t g = hipGetDevice(&dev);
  if (g != hipSuccess) HM_FAIL(HM_ERR_HIP, "hm_serving_field_noalloc: hipGetDevice: %s", hipGetErrorString(g));
  if (dev < 0 || dev >= FS_MAX_DEVICES) HM_FAIL(HM_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, FS_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      HM_FAIL(HM_ERR_HIP, "hm_serving_field_noalloc: the workspace's address on device %d: %s", dev,
              hipGetErrorString(e));
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return HM_OK;
}

}  // namespace hm

using namespace hm;

extern "C" int hm_serving_normal(const int64_t* seed, int64_t salt, int64_t n, double* out, void* stream) {
  HM_REQUIRE(seed && n >= 0 && (n == 0 || out), "bad normal draw arguments");
  if (n == 0) return HM_OK;
  hipLaunchKernelGGL(k_normal, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, seed, salt, n, out);
  return hip_check("hm_serving_normal");
}

extern "C" int hm_serving_field(const int64_t* seed, int n1, int passes, double* work, const int64_t* lattice_to_dof,
                                double* out, void* stream) {
  HM_REQUIRE(seed && work && out && n1 >= 1 && passes >= 0, "bad field arguments");
  return field_into(seed, n1, passes, work, lattice_to_dof, out, (hipStream_t)stream, "hm_serving_field");
}

extern "C" int hm_serving_field_noalloc(const int64_t* seed, int n1, int passes, const int64_t* lattice_to_dof,
                                        double* out, void* stream) {
  HM_REQUIRE(seed && out && n1 >= 1 && passes >= 0, "bad field arguments");
  if (n1 > FS_NOALLOC_MAX_N1)
    HM_FAIL(HM_ERR_ARG, "hm_serving_field_noalloc: a lattice of %d points a direction exceeds its workspace of %d",
            n1, FS_NOALLOC_MAX_N1);
  double* work = nullptr;
  const int rc = scratch_of_device(&work);
  if (rc != HM_OK) return rc;
  return field_into(seed, n1, passes, work, lattice_to_dof, out, (hipStream_t)stream, "hm_serving_field_noalloc");
}
