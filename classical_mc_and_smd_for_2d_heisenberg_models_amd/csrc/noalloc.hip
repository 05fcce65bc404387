// Energies of caller-owned device configurations without any allocation:
// the variant of hsmd_energies (get_energyValue, CF/subrouts.cpp:5181-5312)
// for callers that cannot allocate device memory.
//
// The bond list comes from an MC engine (device-resident since its
// creation); inputs and outputs are the caller's device buffers; the
// per-block partial sums of the two-stage reduction live in the file-scope
// array `scratch` below, which every GPU the library runs on holds until the
// process ends.  Host code never names `scratch` in a kernel argument or a
// copy: on each device it uses the address hipGetSymbolAddress returns for
// that device (each GPU holds its own copy of the array).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "device_common.h"
#include "hsmd_internal.h"

namespace hsmd {

struct MCEngine;
McBondsDev mc_bonds_dev(const MCEngine &e);

// Partial sums of the reduction: one per (configuration, block).
constexpr int kNoAllocPartials = 1 << 18;   // 2 MiB of FP64
constexpr int kNoAllocBonds = 4096;          // bonds per block
constexpr int kNoAllocThreads = 256;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocPartials];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static double *scratch_on_current_device() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw Error(HSMD_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p)
      throw Error(HSMD_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
    g_scratch_ptr[dev] = (double *)p;
  }
  return g_scratch_ptr[dev];
}

// Stage 1: block g of configuration r sums bonds [g B, (g + 1) B) and the
// field term of sites [g Ns, (g + 1) Ns) into part[r * G + g].
__global__ __launch_bounds__(kNoAllocThreads) void noalloc_energy_partial(
    const McBondsDev b, int G, int site_chunk, const double *spins, double *part) {
  __shared__ double red[16];
  const int g = blockIdx.x, r = blockIdx.y;
  const double *S = spins + (size_t)r * b.N * 3;
  const int b0 = g * kNoAllocBonds, b1 = min(b.bd.Nb, b0 + kNoAllocBonds);
  double e = 0.0;
  for (int k = b0 + threadIdx.x; k < b1; k += blockDim.x) {
    const int i = b.perm[b.bi[k]], j = b.perm[b.bj[k]], c = b.bc[k];
    const double *si = S + 3 * (size_t)i, *sj = S + 3 * (size_t)j;
    e += b.bd.jxy[c] * (si[0] * sj[0] + si[1] * sj[1]) + b.bd.jz[c] * (si[2] * sj[2]);
  }
  if (b.h_on) {
    const int s0 = g * site_chunk, s1 = min(b.N, s0 + site_chunk);
    double mz = 0.0;
    for (int i = s0 + threadIdx.x; i < s1; i += blockDim.x) mz += S[3 * (size_t)i + 2];
    e -= b.h * mz;
  }
  double v[1] = {e};
  block_sum<1>(v, red);
  if (threadIdx.x == 0) part[(size_t)r * G + g] = v[0];
}

// Stage 2: E[r] = sum of the G partials of configuration r, in block order.
__global__ void noalloc_energy_finish(int R, int G, const double *part, double *E) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double e = 0.0;
  for (int g = 0; g < G; g++) e += part[(size_t)r * G + g];
  E[r] = e;
}

void mc_config_energies_noalloc(const MCEngine &e, int R, const double *spins_dev,
                                double *energies_dev, hipStream_t st) {
  if (R < 1) throw Error(HSMD_ERR_ARG, "n_configs must be >= 1");
  const McBondsDev b = mc_bonds_dev(e);
  const int G = std::max(1, (b.bd.Nb + kNoAllocBonds - 1) / kNoAllocBonds);
  if ((long long)R * G > kNoAllocPartials)
    throw Error(HSMD_ERR_ARG, "n_configs x ceil(n_bonds / 4096) = " + std::to_string((long long)R * G) +
                                  " exceeds the scratch buffer's " + std::to_string(kNoAllocPartials) +
                                  " partial sums");
  double *part = scratch_on_current_device();
  const int site_chunk = (b.N + G - 1) / G;
  hipLaunchKernelGGL(noalloc_energy_partial, dim3(G, R), dim3(kNoAllocThreads), 0, st, b, G,
                     site_chunk, spins_dev, part);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(noalloc_energy_finish, dim3((R + 255) / 256), dim3(256), 0, st, R, G, part,
                     energies_dev);
  HIP_CHECK(hipGetLastError());
}

}  // namespace hsmd
