// axb_energy.hip -- pair energy over a fixed neighbour CRS and its gradient.
//
// No reference counterpart; the pattern is the reference example's
// Lennard-Jones loop over a radius CRS
// (examples/molecular_dynamics/example_molecular_dynamics.cpp:128-194).
// phi(r) = 4 (r^-12 - r^-6) - phi(cutoff);  E = 1/2 sum_i sum_{j in row i,
// j != i} phi(r_ij);  dE/dx_i = sum_{j in row i, j != i} phi'(r)/r (x_i - x_j)
// (rows of a radius query are symmetric).  Pair terms are float, every sum
// is double; the block partials are reduced in a fixed order, so E is
// bit-reproducible run to run.  One lane per row: forward and backward read
// the row once (the fused kernel costs one gather of x_j per pair).
#include "axb_internal.hpp"

#include <mutex>

namespace axb
{
namespace
{

AXB_D double wave_sum_f64(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o, 64);
  return v;
}

// A group of ROW_LANES consecutive lanes per row: lane l of the group takes
// the row's pairs l, l + ROW_LANES, ...; the group then adds its lanes'
// sums by xor shuffles.  The group reads its row's indices as contiguous
// 32-byte pieces and gathers the neighbours of one point at a time, so the
// gathers of a wavefront stay in a small region (its 8 rows are
// consecutive); one lane per row spread each step's loads over 64 rows.
// A block takes ROWS_PER_BLOCK rows at a time, in a grid-stride loop over
// a grid of at most ENERGY_SCRATCH_BLOCKS blocks (the partial sums).
constexpr int ROW_LANES = 8;
constexpr int ROWS_PER_BLOCK = 256 / ROW_LANES;

template <bool WANT_G>
__global__ __launch_bounds__(256) void k_energy(const float *__restrict__ xyz, int64_t n,
                                                const int64_t *__restrict__ offsets,
                                                const int32_t *__restrict__ indices, float phic,
                                                const uint32_t *__restrict__ order,
                                                double *__restrict__ partials,
                                                float *__restrict__ grad)
{
  // rows of neighbouring points share neighbours: rows are taken in Morton
  // order of their points (`order`, when given), one contiguous range of
  // blocks per XCD (its L2), so the x_j gathered by a wavefront and by its
  // XCD overlap
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int sub = (int)(threadIdx.x % ROW_LANES);
  double e = 0.0;
  for (int64_t base = (int64_t)blk * ROWS_PER_BLOCK; base < n;
       base += (int64_t)gridDim.x * ROWS_PER_BLOCK)
  {
    const int64_t t = base + threadIdx.x / ROW_LANES;
    double er = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    int64_t i = 0;
    if (t < n)
    {
      i = order ? (int64_t)order[t] : t;
      const float xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
      const int64_t end = offsets[i + 1];
      for (int64_t p = offsets[i] + sub; p < end; p += ROW_LANES)
      {
        const int64_t j = indices[p];
        if (j == i)
          continue;
        const float dx = xi - xyz[3 * j], dy = yi - xyz[3 * j + 1], dz = zi - xyz[3 * j + 2];
        float r2 = dx * dx;
        r2 = r2 + dy * dy;
        r2 = r2 + dz * dz;
        const float ir2 = 1.0f / r2;
        const float ir6 = ir2 * ir2 * ir2;
        er += (double)(4.0f * ir6 * (ir6 - 1.0f) - phic);
        if (WANT_G)
        {
          const float f = 24.0f * ir2 * ir6 * (1.0f - 2.0f * ir6);
          gx += (double)(f * dx);
          gy += (double)(f * dy);
          gz += (double)(f * dz);
        }
      }
    }
#pragma unroll
    for (int o = ROW_LANES / 2; o > 0; o >>= 1)
    {
      er += __shfl_xor(er, o, 64);
      if (WANT_G)
      {
        gx += __shfl_xor(gx, o, 64);
        gy += __shfl_xor(gy, o, 64);
        gz += __shfl_xor(gz, o, 64);
      }
    }
    if (t < n && sub == 0)
    {
      e += 0.5 * er;
      if (WANT_G)
      {
        grad[3 * i] = (float)gx;
        grad[3 * i + 1] = (float)gy;
        grad[3 * i + 2] = (float)gz;
      }
    }
  }
  __shared__ double s[4];
  e = wave_sum_f64(e);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    s[wave] = e;
  __syncthreads();
  if (threadIdx.x == 0)
    partials[blk] = ((s[0] + s[1]) + s[2]) + s[3];
}

// Blocks of k_energy for n rows (and partial sums): at most
// ENERGY_SCRATCH_BLOCKS, the grid-stride loop takes the rest.
int64_t energy_blocks(int64_t n)
{
  const int64_t need = div_up(n > 0 ? n : 1, ROWS_PER_BLOCK);
  return need < ENERGY_SCRATCH_BLOCKS ? need : ENERGY_SCRATCH_BLOCKS;
}

__global__ __launch_bounds__(256) void k_reduce_f64(const double *__restrict__ partials,
                                                    int64_t m, double *__restrict__ out)
{
  double v = 0.0;
  for (int64_t j = threadIdx.x; j < m; j += 256)
    v += partials[j];
  __shared__ double s[4];
  v = wave_sum_f64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    s[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    *out = ((s[0] + s[1]) + s[2]) + s[3];
}

// FP64 central differences of the row energy of each listed atom (the only
// terms of E that depend on x_i when rows are symmetric).
__global__ __launch_bounds__(256) void k_energy_fd(const float *__restrict__ xyz,
                                                   const int64_t *__restrict__ offsets,
                                                   const int32_t *__restrict__ indices,
                                                   double phic, const int32_t *__restrict__ atoms,
                                                   int64_t m, double h, double *__restrict__ fd)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * m)
    return;
  const int64_t a = t / 3;
  const int d = (int)(t % 3);
  const int64_t i = atoms[a];
  double e2[2];
  for (int sgn = 0; sgn < 2; ++sgn)
  {
    double x[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    x[d] += sgn == 0 ? h : -h;
    double e = 0.0;
    for (int64_t p = offsets[i]; p < offsets[i + 1]; ++p)
    {
      const int64_t j = indices[p];
      if (j == i)
        continue;
      const double dx = x[0] - (double)xyz[3 * j], dy = x[1] - (double)xyz[3 * j + 1],
                   dz = x[2] - (double)xyz[3 * j + 2];
      const double ir2 = 1.0 / (dx * dx + dy * dy + dz * dz);
      const double ir6 = ir2 * ir2 * ir2;
      e += 4.0 * ir6 * (ir6 - 1.0) - phic;
    }
    e2[sgn] = e;
  }
  fd[t] = (e2[0] - e2[1]) / (2.0 * h);
}

// Scratch of the allocation-free entry point: the per-block partial sums
// (one double per 256 rows).  Every GPU the library runs on holds its own
// copy until the process ends.  Host code never passes `energy_scratch`
// itself (in host code the name is a stand-in, not a device address); it
// uses the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double energy_scratch[ENERGY_SCRATCH_BLOCKS];

hipError_t launch_energy(const float *xyz, int64_t n, const int64_t *offsets,
                         const int32_t *indices, float phic, const uint32_t *order,
                         double *partials, int64_t nb, double *energy, float *grad,
                         hipStream_t s, bool profiled)
{
  if (profiled)
  {
    if (grad)
      AXB_LAUNCH(k_energy<true>, dim3((unsigned)nb), dim3(256), 0, s, xyz, n, offsets,
                 indices, phic, order, partials, grad);
    else
      AXB_LAUNCH(k_energy<false>, dim3((unsigned)nb), dim3(256), 0, s, xyz, n, offsets,
                 indices, phic, order, partials, grad);
    if (energy)
      AXB_LAUNCH(k_reduce_f64, dim3(1), dim3(256), 0, s, partials, nb, energy);
  }
  else
  {
    if (grad)
      hipLaunchKernelGGL(k_energy<true>, dim3((unsigned)nb), dim3(256), 0, s, xyz, n, offsets,
                         indices, phic, order, partials, grad);
    else
      hipLaunchKernelGGL(k_energy<false>, dim3((unsigned)nb), dim3(256), 0, s, xyz, n,
                         offsets, indices, phic, order, partials, grad);
    if (energy)
      hipLaunchKernelGGL(k_reduce_f64, dim3(1), dim3(256), 0, s, partials, nb, energy);
  }
  return hipGetLastError();
}

float shifted_phi_cutoff(float cutoff)
{
  const double irc2 = 1.0 / ((double)cutoff * (double)cutoff);
  const double irc6 = irc2 * irc2 * irc2;
  return (float)(4.0 * irc6 * (irc6 - 1.0));
}

} // namespace

hipError_t pair_energy(const float *xyz, int64_t n, const int64_t *offsets,
                       const int32_t *indices, float cutoff, const uint32_t *order,
                       double *energy, float *grad, hipStream_t s)
{
  const float phic = shifted_phi_cutoff(cutoff);
  const int64_t nb = energy_blocks(n);
  double *partials = (double *)dalloc(nb * sizeof(double), s);
  if (!partials)
    return hipErrorOutOfMemory;
  const hipError_t e =
      launch_energy(xyz, n, offsets, indices, phic, order, partials, nb, energy, grad, s, true);
  dfree(partials, s);
  return e;
}

// Device address of this GPU's energy_scratch, looked up once per device
// (with that device current) and kept.
static hipError_t energy_scratch_ptr(double **out)
{
  static std::mutex m;
  static double *ptrs[MAX_DEVICES] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return e;
  if (dev < 0 || dev >= MAX_DEVICES)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(m);
  if (!ptrs[dev])
  {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(energy_scratch));
    if (e != hipSuccess)
      return e;
    ptrs[dev] = (double *)p;
  }
  *out = ptrs[dev];
  return hipSuccess;
}

hipError_t pair_energy_noalloc(const float *xyz, int64_t n, const int64_t *offsets,
                               const int32_t *indices, float cutoff, double *energy,
                               float *grad, hipStream_t s)
{
  if (n > ENERGY_SCRATCH_BLOCKS * 256)
    return hipErrorInvalidValue;
  const int64_t nb = energy_blocks(n);
  double *partials = nullptr;
  const hipError_t e = energy_scratch_ptr(&partials);
  if (e != hipSuccess)
    return e;
  // every block writes its own partial, so nothing needs clearing first
  return launch_energy(xyz, n, offsets, indices, shifted_phi_cutoff(cutoff), nullptr, partials,
                       nb, energy, grad, s, false);
}

hipError_t pair_energy_fd(const float *xyz, const int64_t *offsets, const int32_t *indices,
                          float cutoff, const int32_t *atoms, int64_t m, double h, double *fd,
                          hipStream_t s)
{
  if (m <= 0)
    return hipSuccess;
  const double irc2 = 1.0 / ((double)cutoff * (double)cutoff);
  const double irc6 = irc2 * irc2 * irc2;
  const double phic = 4.0 * irc6 * (irc6 - 1.0);
  AXB_LAUNCH(k_energy_fd, dim3((unsigned)div_up(3 * m, 256)), dim3(256), 0, s, xyz,
                     offsets, indices, phic, atoms, m, h, fd);
  return hipGetLastError();
}

} // namespace axb
