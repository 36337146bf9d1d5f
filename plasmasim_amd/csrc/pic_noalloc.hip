// pic_diagnose_device: the diagnostics of pic_diagnose without any
// allocation, for callers that cannot allocate (code under HIP graph
// capture, tight loops that forbid hipMalloc).
//
// pic_diagnose keeps its workgroup partials in a buffer it allocates with
// hipMalloc on first use and reduces them on the host.  This entry point
// writes the same partials into a static device array, reduces them on the
// device in the same order (quantity by quantity over the workgroups) and
// writes the pic_diag record to caller-provided device memory, all on the
// context's stream, without synchronising.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include <mutex>

#include "pic_internal.h"

namespace pic {
namespace {

constexpr int kMaxDevices = 64;
// the partials of both forms (a layout fixed by the workgroup counts, not by
// the grid or the particle counts) and the live counts of the two species
constexpr int64_t kScratchDoubles = kDiagPartials + 2;
static_assert(kScratchDoubles * sizeof(double) <= 768 * 1024, "scratch stays below 768 KiB");

// One copy per GPU, held until the process ends.  Host code never names
// this array in a launch or a copy: the array's device address on each GPU
// comes from hipGetSymbolAddress (noalloc_init, at pic_create with that
// device current) and is kept in scratch_ptr[device].
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex scratch_mu;
double *scratch_ptr[kMaxDevices];

__global__ void k_clear_counts(unsigned long long *c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0ull;
}

// One workgroup: thread q sums quantity q over the workgroup partials in
// workgroup order, as the host reduction of pic_diagnose does (field
// quantity 5 by max), then thread 0 assembles the record.
__global__ __launch_bounds__(64) void k_diag_finish(const double *__restrict__ pf,
                                                    const unsigned long long *counts,
                                                    pic_diag *out) {
  constexpr int NQ = kDiagNQF + 2 * kDiagNQP;
  static_assert(NQ <= 64, "one thread per quantity");
  __shared__ double s[NQ];
  const int t = threadIdx.x;
  if (t < kDiagNQF) {
    double v = 0.0;
    for (int b = 0; b < kDiagNBF; ++b) {
      const double x = pf[(int64_t)b * kDiagNQF + t];
      v = t == 5 ? fmax(v, x) : v + x;
    }
    s[t] = v;
  } else if (t < NQ) {
    const int sp = (t - kDiagNQF) / kDiagNQP, q = (t - kDiagNQF) % kDiagNQP;
    const double *pp =
        pf + (int64_t)kDiagNBF * kDiagNQF + (int64_t)sp * kDiagNBP * kDiagNQP;
    double v = 0.0;
    for (int b = 0; b < kDiagNBP; ++b) v += pp[(int64_t)b * kDiagNQP + q];
    s[t] = v;
  }
  __syncthreads();
  if (t != 0) return;
  pic_diag d;
  memset(&d, 0, sizeof d);
  d.field_energy_e = s[0];
  d.field_energy_b = s[1];
  d.field_energy_b1 = s[2];
  d.charge = s[3];
  d.gauss_l2 = sqrt(s[4]);
  d.gauss_max = s[5];
  d.energy_box[0] = s[7];
  d.energy_box[1] = s[8];
  for (int sp = 0; sp < 2; ++sp) {
    const double *p = s + kDiagNQF + sp * kDiagNQP;
    d.kinetic[sp] = p[0];
    d.kinetic_nr[sp] = p[1];
    for (int j = 0; j < 3; ++j) d.momentum[sp][j] = p[2 + j];
    for (int j = 0; j < 3; ++j) d.force[j] += p[5 + j];
    const int b0 = sp == PIC_ION ? 3 : 9;
    for (int j = 0; j < 6; ++j) d.force[b0 + j] = p[8 + j];
    d.energy_box[2 + sp] = p[20];
    d.count[sp] = (int64_t)counts[sp];
  }
  *out = d;
}

}  // namespace

int noalloc_init(pic_ctx *ctx) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    fprintf(stderr, "pic_create: hipGetDevice: %s\n", hipGetErrorString(e));
    return PIC_E_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    fprintf(stderr, "pic_create: device %d beyond the %d supported\n", dev, kMaxDevices);
    return PIC_E_HIP;
  }
  ctx->device = dev;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "pic_create: hipGetSymbolAddress(scratch): %s\n",
              hipGetErrorString(e));
      return PIC_E_HIP;
    }
    scratch_ptr[dev] = (double *)p;
  }
  ctx->noalloc_scratch = scratch_ptr[dev];
  return 0;
}

}  // namespace pic

using namespace pic;

// No reference counterpart: the allocation-free variant of pic_diagnose
// (thrust_f :4940-5122 and the Gauss residual :8345-8348).
extern "C" int pic_diagnose_device(pic_ctx *ctx, pic_diag *dev_out) {
  if (!ctx || !dev_out) return PIC_E_ARG;
  double *pf = ctx->noalloc_scratch;
  if (!pf) return set_error(ctx, PIC_E_STATE, "diagnose_device: no scratch on this device");
  // The partial layout does not grow with the grid or the particle counts
  // (kDiagNBF + 2 kDiagNBP workgroup partials), so every context fits the
  // scratch; this guards the constant should the layout change.
  if (kDiagPartials + 2 > kScratchDoubles)
    return set_error(ctx, PIC_E_CAPACITY,
                     "diagnose_device: %lld partials exceed the scratch of %lld doubles",
                     (long long)(kDiagPartials + 2), (long long)kScratchDoubles);
  unsigned long long *counts = (unsigned long long *)(pf + kDiagPartials);
  int rc;
  const bool f64 = ctx->cfg.precision == PIC_FP64;
  if (!f64 && (rc = f32::finish_deposit(ctx))) return rc;
  // the counts accumulate with atomics: cleared by a kernel (not a memset
  // node, so the call can be captured into a graph)
  k_clear_counts<<<1, 64, 0, ctx->stream>>>(counts);
  PIC_LAUNCH_CHECK(ctx);
  rc = f64 ? f64::diagnose_partials(ctx, pf, counts)
           : f32::diagnose_partials(ctx, pf, counts);
  if (rc) return rc;
  k_diag_finish<<<1, 64, 0, ctx->stream>>>(pf, counts, dev_out);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}
