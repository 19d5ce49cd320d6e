// fmi_ingest.hip -- reads in the reference's fastq_batch_t layout
// (bioformats/fastq/fastq_batch.h:24-34: the sequences one after another in
// `seq`, each followed by a NUL, and int data_indices[num_reads + 1] with
// the start of each) turned into the engine's read batch on the device: the
// base codes of encode_bases (bwt_commons.c:99-102, alphabet "ACGT"; any
// other byte is N, code 4), without the NULs, and uint64 offsets; and,
// optionally, the codes packed 2 bits per base with a 1-bit N mask.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_kernels.h"

namespace fmi {

namespace {

__device__ __forceinline__ uint8_t encode_base(uint8_t c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}

// One thread per byte seq[data_indices[0] + p], p < nbytes = data_indices
// [nreads] - data_indices[0] (data_indices index seq, as the reference's
// batch->seq[batch->data_indices[r]]): the read it belongs to by binary
// search of
// data_indices; its NUL terminator is dropped, every other byte goes to
// codes[base_off + data_indices[r] - data_indices[0] - r + j].  Threads
// p <= nreads also write offsets[read_off + p] = data_indices[p] -
// data_indices[0] - p (from 0: a caller placing several batches side by
// side adds each one's base_off afterwards).  Indices that are not
// increasing by at least one per read give wrong codes but no write outside
// codes[base_off .. base_off + nbytes - nreads).
__global__ __launch_bounds__(256) void fastq_encode_kernel(const char *seq, const int32_t *di,
                                                           uint32_t nreads, uint64_t nbytes,
                                                           uint64_t work, uint64_t base_off,
                                                           uint64_t read_off, uint8_t *codes,
                                                           uint64_t *offsets) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const int64_t d0 = di[0];
  const uint64_t ncodes = nbytes >= nreads ? nbytes - nreads : 0;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < work; p += stride) {
    if (p <= nreads) {
      // each earlier read left one NUL behind
      offsets[read_off + p] = (uint64_t)((int64_t)di[p] - d0 - (int64_t)p);
    }
    if (p >= nbytes || nreads == 0) continue;
    // the read holding byte d0 + p: the last r with di[r] <= d0 + p
    const int64_t at = d0 + (int64_t)p;
    uint32_t lo = 0, hi = nreads;            // di[lo] <= at < di[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if ((int64_t)di[mid] <= at) lo = mid;
      else hi = mid;
    }
    if (at == (int64_t)di[lo + 1] - 1) continue;          // its NUL
    const int64_t o = at - d0 - (int64_t)lo;
    if (o < 0 || (uint64_t)o >= ncodes) continue;
    codes[base_off + (uint64_t)o] = encode_base((uint8_t)seq[at]);
  }
}

// 2-bit packing of nbases codes: thread t packs bases 16t .. 16t + 15 into
// packed[t] (base i at bits 2(i % 16)) and their N flags into nmask's bits
// (base i at bit i % 16 of the 16-bit word t); an N packs as 0.
__global__ __launch_bounds__(256) void pack_codes_kernel(const uint8_t *codes, uint64_t nbases,
                                                         uint32_t *packed, uint16_t *nmask) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t words = (nbases + 15) / 16;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < words; t += stride) {
    uint32_t w = 0;
    uint32_t m = 0;
    for (uint32_t j = 0; j < 16; ++j) {
      const uint64_t i = 16 * t + j;
      if (i >= nbases) break;
      const uint32_t c = codes[i];
      if (c > 3u) m |= 1u << j;
      else w |= c << (2 * j);
    }
    packed[t] = w;
    if (nmask) nmask[t] = (uint16_t)m;
  }
}

}  // namespace

hipError_t launch_fastq_encode(const char *seq, const int32_t *data_indices, uint32_t nreads,
                               uint64_t nbytes, uint64_t base_off, uint64_t read_off,
                               uint8_t *codes, uint64_t *offsets, hipStream_t s) {
  const uint64_t work = nbytes > nreads + 1 ? nbytes : nreads + 1;
  uint64_t grid = (work + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(fastq_encode_kernel, dim3((uint32_t)grid), dim3(256), 0, s, seq,
                     data_indices, nreads, nbytes, work, base_off, read_off, codes, offsets);
  return hipGetLastError();
}

hipError_t launch_pack_codes(const uint8_t *codes, uint64_t nbases, uint32_t *packed,
                             uint16_t *nmask, hipStream_t s) {
  const uint64_t words = (nbases + 15) / 16;
  if (words == 0) return hipSuccess;
  uint64_t grid = (words + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(pack_codes_kernel, dim3((uint32_t)grid), dim3(256), 0, s, codes, nbases,
                     packed, nmask);
  return hipGetLastError();
}

}  // namespace fmi
