// fmi_sa_gpu.hip -- index construction on the GPU (fmi_index_build_gpu).
//
// The reference builds offline on the host (bwt_generate_index_files,
// bwt.c:744-819, dbwt + an n-step serial LF walk for the samples), limited to
// texts below 2^31.  At 3.1 Gbp that is hours; here the suffix array of each
// direction is sorted on the GPU by prefix doubling, and the occurrence
// blocks and samples are written from it by GPU kernels straight into the
// index's device allocation.
//
// Suffix sorting (suffixes 0..n-1; the '$' suffix n is row 0 of the BWT):
//   round 0   key = the first 21 symbols, 3 bits each (code + 1; 0 past the
//             end, so '$' sorts first); radix sort (key, position).
//   rank[i]   = SA slot of the first member of i's group (equal prefixes).
//   round r   only suffixes in groups of size >= 2 take part: key =
//             (rank[i], rank[i + h] + 1, or 0 past the end), radix sort,
//             write back into their own slots, new group heads, h *= 2.
//             Stops when every suffix is alone in its group.
// A suffix within h of the end is alone after the round that sees '$' in its
// key, so rank[i + h] is only read for i + h < n.
#include <hip/hip_runtime.h>

#include <cstring>   // rocprim's texture_cache_iterator uses memset

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_index.h"
#include "fmi_internal.h"

namespace fmi {

namespace {

constexpr int kRound0Syms = 21;

__global__ void round0_keys(const uint8_t *T, uint64_t n, uint64_t *keys, uint32_t *vals) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = 0;
#pragma unroll
    for (int t = 0; t < kRound0Syms; ++t) {
      const uint64_t c = (i + t < n) ? (uint64_t)T[i + t] + 1u : 0u;
      k = (k << 3) | c;
    }
    keys[i] = k;
    vals[i] = (uint32_t)i;
  }
}

// hp[t] = slot of t if t starts a group (key differs from t-1), else 0.
__global__ void head_slots(const uint64_t *keys, const uint32_t *slot_of, uint64_t m, uint32_t *hp) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = slot_of ? slot_of[t] : (uint32_t)t;
    hp[t] = (t == 0 || keys[t] != keys[t - 1]) ? s : 0u;
  }
}

// rank[suffix] = group start slot; flag[t] = 1 when t's group has >= 2 members.
__global__ void assign_ranks(const uint32_t *sorted_pos, const uint32_t *gs, const uint32_t *slot_of,
                             uint64_t m, uint32_t *rank, uint8_t *flag) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = slot_of ? slot_of[t] : (uint32_t)t;
    rank[sorted_pos[t]] = gs[t];
    const bool head = gs[t] == s;
    const bool next_head = (t + 1 == m) ||
                           (gs[t + 1] == (slot_of ? slot_of[t + 1] : (uint32_t)(t + 1)));
    flag[t] = (head && next_head) ? 0u : 1u;
  }
}

__global__ void round_keys(const uint32_t *U, uint64_t m, const uint32_t *SA, const uint32_t *rank,
                           uint64_t n, uint64_t h, uint64_t *keys, uint32_t *vals) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = SA[U[t]];
    const uint64_t r2 = ((uint64_t)i + h < n) ? (uint64_t)rank[(uint64_t)i + h] + 1u : 0u;
    keys[t] = ((uint64_t)rank[i] << 32) | r2;
    vals[t] = i;
  }
}

__global__ void write_back(const uint32_t *U, const uint32_t *vals, uint64_t m, uint32_t *SA) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (uint64_t)gridDim.x * blockDim.x)
    SA[U[t]] = vals[t];
}

__global__ void iota_u32(uint32_t *p, uint64_t m) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (uint64_t)gridDim.x * blockDim.x)
    p[t] = (uint32_t)t;
}

__global__ void reverse_text(const uint8_t *T, uint64_t n, uint8_t *R) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    R[i] = T[n - 1 - i];
}

// One wavefront per 192-row block: three ballots per plane build the plane
// words; per-block symbol counts go to bc[b * nb + j] for the scan.
__global__ __launch_bounds__(64) void blocks_kernel(const uint8_t *T, const uint32_t *SA, uint64_t n,
                                                   uint64_t nb, OccBlock *blk, uint32_t *bc,
                                                   uint32_t *S, uint32_t *R, uint32_t ratio,
                                                   unsigned long long *dollar) {
  const uint32_t lane = threadIdx.x;
  for (uint64_t j = blockIdx.x; j < nb; j += gridDim.x) {
    uint32_t c[4] = {0, 0, 0, 0};
    uint64_t lo[3], hi[3];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      const uint64_t r = j * kBlockRows + (uint64_t)w * 64 + lane;
      const bool valid = r <= n;
      uint64_t p = 0;
      if (valid) p = (r == 0) ? n : (uint64_t)SA[r - 1];
      const bool sym = valid && p > 0;
      const uint32_t code = sym ? T[p - 1] : 0u;
      if (valid && p == 0) *dollar = r;
      if (valid && r % ratio == 0) S[r / ratio] = (uint32_t)p;
      if (valid && p % ratio == 0) R[p / ratio] = (uint32_t)r;
      lo[w] = __ballot(sym && (code & 1u));
      hi[w] = __ballot(sym && (code & 2u));
      const uint64_t any = __ballot(sym);
      const uint32_t n3 = (uint32_t)__popcll(lo[w] & hi[w]);
      const uint32_t n2 = (uint32_t)__popcll(hi[w]) - n3;
      const uint32_t n1 = (uint32_t)__popcll(lo[w]) - n3;
      c[3] += n3; c[2] += n2; c[1] += n1;
      c[0] += (uint32_t)__popcll(any) - n1 - n2 - n3;
    }
    if (lane == 0) {
#pragma unroll
      for (int w = 0; w < 3; ++w) { blk[j].lo[w] = lo[w]; blk[j].hi[w] = hi[w]; }
#pragma unroll
      for (int b = 0; b < 4; ++b) bc[(uint64_t)b * nb + j] = c[b];
    }
  }
}

__global__ void block_counts(const uint32_t *ex, uint64_t nb, OccBlock *blk) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nb;
       j += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int b = 0; b < 4; ++b) blk[j].cnt[b] = ex[(uint64_t)b * nb + j];
  }
}

inline dim3 grid_for(uint64_t m) {
  uint64_t g = (m + 255) / 256;
  if (g > 65536ull * 16) g = 65536ull * 16;
  if (g == 0) g = 1;
  return dim3((uint32_t)g);
}

#define GPU_TRY(expr)                                     \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) {                               \
      *err = std::string(#expr) + ": " + hipGetErrorString(e_); \
      return FMI_EDEVICE;                                 \
    }                                                     \
  } while (0)

struct Scratch {
  uint64_t *k0 = nullptr, *k1 = nullptr;   // keys (double buffer)
  uint32_t *v0 = nullptr;                  // values (other half of the double buffer is SA)
  uint32_t *rank = nullptr;
  uint32_t *U = nullptr;                   // unresolved slots
  uint32_t *hp = nullptr, *gs = nullptr;   // head slots, group starts
  uint8_t *flag = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  uint64_t *count = nullptr;               // device counter for select
};

// Sorts the n non-empty suffixes of T (device) into SA (device, n entries).
int gpu_suffix_array(const uint8_t *T, uint64_t n, uint32_t *SA, Scratch &w, hipStream_t s,
                     std::string *err) {
  if (n == 0) return FMI_OK;
  // round 0
  round0_keys<<<grid_for(n), 256, 0, s>>>(T, n, w.k0, w.v0);
  {
    rocprim::double_buffer<uint64_t> kb(w.k0, w.k1);
    rocprim::double_buffer<uint32_t> vb(w.v0, SA);
    size_t need = 0;
    GPU_TRY(rocprim::radix_sort_pairs(nullptr, need, kb, vb, n, 0, 3 * kRound0Syms, s));
    if (need > w.tmp_bytes) { *err = "sort scratch too small"; return FMI_ENOMEM; }
    GPU_TRY(rocprim::radix_sort_pairs(w.tmp, need, kb, vb, n, 0, 3 * kRound0Syms, s));
    if (vb.current() != SA)
      GPU_TRY(hipMemcpyAsync(SA, vb.current(), n * 4, hipMemcpyDeviceToDevice, s));
    uint64_t *sk = kb.current();
    head_slots<<<grid_for(n), 256, 0, s>>>(sk, nullptr, n, w.hp);
  }
  size_t need = 0;
  GPU_TRY(rocprim::inclusive_scan(nullptr, need, w.hp, w.gs, n, rocprim::maximum<uint32_t>(), s));
  if (need > w.tmp_bytes) { *err = "scan scratch too small"; return FMI_ENOMEM; }
  GPU_TRY(rocprim::inclusive_scan(w.tmp, need, w.hp, w.gs, n, rocprim::maximum<uint32_t>(), s));
  assign_ranks<<<grid_for(n), 256, 0, s>>>(SA, w.gs, nullptr, n, w.rank, w.flag);
  // unresolved slots: select t with flag[t]
  iota_u32<<<grid_for(n), 256, 0, s>>>(w.hp, n);
  GPU_TRY(rocprim::select(nullptr, need, w.hp, w.flag, w.U, w.count, n, s));
  if (need > w.tmp_bytes) { *err = "select scratch too small"; return FMI_ENOMEM; }
  GPU_TRY(rocprim::select(w.tmp, need, w.hp, w.flag, w.U, w.count, n, s));
  uint64_t m = 0;
  GPU_TRY(hipMemcpyAsync(&m, w.count, 8, hipMemcpyDeviceToHost, s));
  GPU_TRY(hipStreamSynchronize(s));

  uint64_t h = kRound0Syms;
  for (int round = 1; m > 0; ++round) {
    if (round > 40 || h > 2 * n + 64) { *err = "suffix sorting did not converge"; return FMI_EDEVICE; }
    round_keys<<<grid_for(m), 256, 0, s>>>(w.U, m, SA, w.rank, n, h, w.k0, w.v0);
    rocprim::double_buffer<uint64_t> kb(w.k0, w.k1);
    rocprim::double_buffer<uint32_t> vb(w.v0, w.gs);   // gs is free here
    GPU_TRY(rocprim::radix_sort_pairs(nullptr, need, kb, vb, m, 0, 64, s));
    if (need > w.tmp_bytes) { *err = "sort scratch too small"; return FMI_ENOMEM; }
    GPU_TRY(rocprim::radix_sort_pairs(w.tmp, need, kb, vb, m, 0, 64, s));
    uint32_t *sorted = vb.current();
    write_back<<<grid_for(m), 256, 0, s>>>(w.U, sorted, m, SA);
    uint32_t *gsr = (sorted == w.gs) ? w.v0 : w.gs;   // the buffer not holding `sorted`
    head_slots<<<grid_for(m), 256, 0, s>>>(kb.current(), w.U, m, w.hp);
    GPU_TRY(rocprim::inclusive_scan(nullptr, need, w.hp, gsr, m, rocprim::maximum<uint32_t>(), s));
    if (need > w.tmp_bytes) { *err = "scan scratch too small"; return FMI_ENOMEM; }
    GPU_TRY(rocprim::inclusive_scan(w.tmp, need, w.hp, gsr, m, rocprim::maximum<uint32_t>(), s));
    assign_ranks<<<grid_for(m), 256, 0, s>>>(sorted, gsr, w.U, m, w.rank, w.flag);
    // next unresolved set (U is read by select and written to hp: no overlap)
    GPU_TRY(rocprim::select(nullptr, need, w.U, w.flag, w.hp, w.count, m, s));
    if (need > w.tmp_bytes) { *err = "select scratch too small"; return FMI_ENOMEM; }
    GPU_TRY(rocprim::select(w.tmp, need, w.U, w.flag, w.hp, w.count, m, s));
    std::swap(w.U, w.hp);
    GPU_TRY(hipMemcpyAsync(&m, w.count, 8, hipMemcpyDeviceToHost, s));
    GPU_TRY(hipStreamSynchronize(s));
    h *= 2;
  }
  return FMI_OK;
}

// Occurrence blocks, C and samples of one direction from its suffix array.
int gpu_fill_direction(const uint8_t *T, uint64_t n, const uint32_t *SA, uint32_t ratio,
                       OccBlock *blk, uint64_t nb, uint32_t *S, uint32_t *R, Scratch &w,
                       HostDir *hd, hipStream_t s, std::string *err) {
  uint32_t *bc = reinterpret_cast<uint32_t *>(w.k0);       // 4 * nb counts
  uint32_t *ex = reinterpret_cast<uint32_t *>(w.k1);       // their exclusive scans
  unsigned long long *dollar = reinterpret_cast<unsigned long long *>(w.count);
  GPU_TRY(hipMemsetAsync(dollar, 0, 8, s));
  uint64_t g = nb < 65536ull * 4 ? nb : 65536ull * 4;
  blocks_kernel<<<dim3((uint32_t)g), 64, 0, s>>>(T, SA, n, nb, blk, bc, S, R, ratio, dollar);
  size_t need = 0;
  for (int b = 0; b < 4; ++b) {
    GPU_TRY(rocprim::exclusive_scan(nullptr, need, bc + (uint64_t)b * nb, ex + (uint64_t)b * nb, 0u, nb,
                                    rocprim::plus<uint32_t>(), s));
    if (need > w.tmp_bytes) { *err = "scan scratch too small"; return FMI_ENOMEM; }
    GPU_TRY(rocprim::exclusive_scan(w.tmp, need, bc + (uint64_t)b * nb, ex + (uint64_t)b * nb, 0u, nb,
                                    rocprim::plus<uint32_t>(), s));
  }
  block_counts<<<grid_for(nb), 256, 0, s>>>(ex, nb, blk);
  // C from the totals: exclusive prefix of the last block plus its own count
  uint32_t last_ex[4], last_bc[4];
  for (int b = 0; b < 4; ++b) {
    GPU_TRY(hipMemcpyAsync(&last_ex[b], ex + (uint64_t)b * nb + nb - 1, 4, hipMemcpyDeviceToHost, s));
    GPU_TRY(hipMemcpyAsync(&last_bc[b], bc + (uint64_t)b * nb + nb - 1, 4, hipMemcpyDeviceToHost, s));
  }
  unsigned long long dl = 0;
  GPU_TRY(hipMemcpyAsync(&dl, dollar, 8, hipMemcpyDeviceToHost, s));
  GPU_TRY(hipStreamSynchronize(s));
  uint64_t acc = 0;
  for (int b = 0; b < 4; ++b) { hd->C[b] = (uint32_t)acc; acc += (uint64_t)last_ex[b] + last_bc[b]; }
  if (acc != n) { *err = "block counts do not add up to the text length"; return FMI_EDEVICE; }
  hd->dollar = dl;
  return FMI_OK;
}

}  // namespace

int build_index_gpu(const uint8_t *text, uint64_t n, uint32_t ratio, fmi_index *ix, std::string *err) {
  if (!text && n) { *err = "text is NULL"; return FMI_EINVAL; }
  if (ratio == 0) { *err = "sa_ratio must be >= 1"; return FMI_EINVAL; }
  if (n == 0) { *err = "empty text (use fmi_index_build)"; return FMI_EINVAL; }
  if (n + 3 >= (uint64_t)UINT32_MAX) { *err = "text longer than 2^32 - 4 symbols"; return FMI_ERANGE; }
  for (uint64_t i = 0; i < n; ++i)
    if (text[i] > 3) { *err = "text holds a code > 3 (replace N before indexing)"; return FMI_EINVAL; }
  int dev = 0;
  GPU_TRY(hipGetDevice(&dev));
  const uint64_t rows = n + 1;
  const uint64_t nb = (n + 2) / kBlockRows + 1;
  const uint64_t nsamp = rows / ratio + (rows % ratio ? 1 : 0);
  const uint64_t samp_bytes = (nsamp * 4 + 255) & ~255ull;
  const uint64_t index_bytes = 2 * (nb * sizeof(OccBlock) + 2 * samp_bytes);

  // scratch: text x2, SA, keys x2, values, rank, U, hp, gs, flag, sort temp
  Scratch w;
  size_t need = 0, t1 = 0, t2 = 0, t3 = 0;
  {
    rocprim::double_buffer<uint64_t> kb(nullptr, nullptr);
    rocprim::double_buffer<uint32_t> vb(nullptr, nullptr);
    if (rocprim::radix_sort_pairs(nullptr, t1, kb, vb, n, 0, 64, (hipStream_t)0) != hipSuccess) t1 = 0;
    uint32_t *np = nullptr;
    if (rocprim::inclusive_scan(nullptr, t2, np, np, n, rocprim::maximum<uint32_t>(), (hipStream_t)0) !=
        hipSuccess) t2 = 0;
    uint8_t *fp = nullptr;
    uint64_t *cp = nullptr;
    if (rocprim::select(nullptr, t3, np, fp, np, cp, n, (hipStream_t)0) != hipSuccess) t3 = 0;
  }
  need = std::max(t1, std::max(t2, t3)) + (1u << 20);
  const uint64_t scratch_bytes = 2 * n + 4 * n /*SA*/ + 16 * n /*keys*/ + 4 * n /*v0*/ + 4 * n /*rank*/ +
                                 4 * n /*U*/ + 4 * n /*hp*/ + 4 * n /*gs*/ + n /*flag*/ + need + 4096;
  size_t free_b = 0, total_b = 0;
  GPU_TRY(hipMemGetInfo(&free_b, &total_b));
  if (scratch_bytes + index_bytes + (1ull << 30) > free_b) {
    *err = "not enough device memory to build this index on the GPU";
    return FMI_ENOMEM;
  }
  char *mem = nullptr, *imem = nullptr;
  GPU_TRY(hipMalloc(&imem, index_bytes));
  hipError_t ae = hipMalloc(&mem, scratch_bytes);
  if (ae != hipSuccess) {
    (void)hipFree(imem);
    *err = std::string("hipMalloc(build scratch): ") + hipGetErrorString(ae);
    return FMI_ENOMEM;
  }
  auto carve = [&](uint64_t bytes) {
    char *p = mem;
    mem += (bytes + 255) & ~255ull;
    return p;
  };
  char *base = mem;
  uint8_t *dT = reinterpret_cast<uint8_t *>(carve(n));
  uint8_t *dTr = reinterpret_cast<uint8_t *>(carve(n));
  uint32_t *dSA = reinterpret_cast<uint32_t *>(carve(4 * n));
  w.k0 = reinterpret_cast<uint64_t *>(carve(8 * n));
  w.k1 = reinterpret_cast<uint64_t *>(carve(8 * n));
  w.v0 = reinterpret_cast<uint32_t *>(carve(4 * n));
  w.rank = reinterpret_cast<uint32_t *>(carve(4 * n));
  w.U = reinterpret_cast<uint32_t *>(carve(4 * n));
  w.hp = reinterpret_cast<uint32_t *>(carve(4 * n));
  w.gs = reinterpret_cast<uint32_t *>(carve(4 * n));
  w.flag = reinterpret_cast<uint8_t *>(carve(n));
  w.count = reinterpret_cast<uint64_t *>(carve(64));
  w.tmp = carve(need);
  w.tmp_bytes = need;

  hipStream_t s = nullptr;
  int rc = FMI_OK;
  std::string e2;
  auto cleanup = [&](int code) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(base);
    if (code != FMI_OK) (void)hipFree(imem);
    return code;
  };
  if (hipMemcpy(dT, text, n, hipMemcpyHostToDevice) != hipSuccess) {
    *err = "copy text";
    return cleanup(FMI_EDEVICE);
  }
  reverse_text<<<grid_for(n), 256, 0, s>>>(dT, n, dTr);

  ix->n = n;
  ix->ratio = ratio;
  DevIndex dv{};
  char *ip = imem;
  for (int d = 0; d < 2; ++d) {
    HostDir &hd = ix->h[d];
    hd.nblocks = nb;
    hd.nsamp = nsamp;
    OccBlock *dblk = reinterpret_cast<OccBlock *>(ip);
    ip += nb * sizeof(OccBlock);
    uint32_t *dS = reinterpret_cast<uint32_t *>(ip);
    ip += samp_bytes;
    uint32_t *dR = reinterpret_cast<uint32_t *>(ip);
    ip += samp_bytes;
    const uint8_t *Td = d ? dTr : dT;
    rc = gpu_suffix_array(Td, n, dSA, w, s, err);
    if (rc == FMI_OK) rc = gpu_fill_direction(Td, n, dSA, ratio, dblk, nb, dS, dR, w, &hd, s, err);
    if (rc != FMI_OK) return cleanup(rc);
    // host copies (file save, BWT export)
    void *hb = nullptr, *hs = nullptr, *hr = nullptr;
    if (posix_memalign(&hb, 64, nb * sizeof(OccBlock)) || posix_memalign(&hs, 64, nsamp * 4) ||
        posix_memalign(&hr, 64, nsamp * 4)) {
      *err = "out of host memory";
      free(hb); free(hs); free(hr);
      return cleanup(FMI_ENOMEM);
    }
    hd.blk = static_cast<OccBlock *>(hb);
    hd.S = static_cast<uint32_t *>(hs);
    hd.R = static_cast<uint32_t *>(hr);
    if (hipMemcpy(hd.blk, dblk, nb * sizeof(OccBlock), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hd.S, dS, nsamp * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hd.R, dR, nsamp * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      *err = "copy index back";
      return cleanup(FMI_EDEVICE);
    }
    dv.d[d].blk = dblk;
    dv.d[d].S = dS;
    dv.d[d].R = dR;
    for (int b = 0; b < 4; ++b) dv.d[d].C[b] = hd.C[b];
    dv.d[d].dollar = (uint32_t)hd.dollar;
    dv.d[d].dollar_blk = (uint32_t)(hd.dollar / kBlockRows);
    dv.d[d].dollar_off = (uint32_t)(hd.dollar % kBlockRows);
  }
  cleanup(FMI_OK);
  dv.n = (uint32_t)n;
  dv.ratio = ratio;
  ix->kstart.assign(1, 0);
  ix->kend.assign(1, n - 1);
  ix->koffset = {0, n};
  void *ko = nullptr, *ks = nullptr;
  if (hipMalloc(&ko, 16) != hipSuccess || hipMalloc(&ks, 8) != hipSuccess) {
    *err = "hipMalloc(karyotype)";
    (void)hipFree(imem);
    return FMI_ENOMEM;
  }
  (void)hipMemcpy(ko, ix->koffset.data(), 16, hipMemcpyHostToDevice);
  (void)hipMemcpy(ks, ix->kstart.data(), 8, hipMemcpyHostToDevice);
  ix->d_mem = imem;
  ix->d_koffset = static_cast<uint64_t *>(ko);
  ix->d_kstart = static_cast<uint64_t *>(ks);
  ix->dev = dv;
  ix->device = dev;
  ix->device_bytes = index_bytes;
  ix->on_device = true;
  return FMI_OK;
}

}  // namespace fmi
