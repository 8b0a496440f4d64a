// LDS bank-rule probe for gfx950 (not part of libsbdx; built by
// `make -C sbd-pub_amd/csrc probe`).  Every lane of every wavefront reads
// 16-byte LDS elements at a fixed per-lane element index, 4096 times; one
// dispatch per index pattern.  Run under
//   rocprofv3 --pmc SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE -- <this>
// to get LDS cycles per wave instruction for each pattern (the program also
// prints the same figure from HIP-event timing).  Patterns: consecutive
// elements; all lanes on one residue mod 16; and, for every 2-dimensional
// subspace S spanned by two lane-index bits, lanes in the same coset of S
// sharing a residue mod 16 (4 lanes per residue) -- 4 cycles when the 4 lanes
// of each coset are served in different cycles, more when not.  Element
// indices stay below 4096 (64 KB of LDS).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

typedef float V4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const V4 lds_v4;

__global__ void __launch_bounds__(256) k_lds_probe(const unsigned *__restrict__ pat, int iters, float *out) {
  extern __shared__ __align__(16) unsigned char lds[];
  for (int i = threadIdx.x; i < 4096; i += 256) reinterpret_cast<V4 *>(lds)[i] = V4((float)i);
  __syncthreads();
  unsigned a[8];
  for (int k = 0; k < 8; k++) a[k] = (pat[threadIdx.x & 63] & 4095u) * 16u;
  float acc = 0.f;
  for (int it = 0; it < iters; it += 8) {
    // empty asm statements (no instruction) keep the 8 addresses opaque, so
    // the 8 reads of an iteration are independent and none is hoisted
#pragma unroll
    for (int k = 0; k < 8; k++) asm volatile("" : "+v"(a[k]));
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const V4 v = *reinterpret_cast<lds_v4 *>((size_t)a[k]);
      acc += v.x + v.w;
    }
  }
  out[blockIdx.x * 256 + threadIdx.x] = acc;
}

// The same for 4-byte reads (ds_read_b32): every lane reads one float at a
// fixed per-lane element index (below 16384, 64 KB of LDS).
__global__ void __launch_bounds__(256) k_lds_probe32(const unsigned *__restrict__ pat, int iters, float *out) {
  extern __shared__ __align__(16) unsigned char lds[];
  typedef __attribute__((address_space(3))) const float lds_f;
  for (int i = threadIdx.x; i < 16384; i += 256) reinterpret_cast<float *>(lds)[i] = (float)i;
  __syncthreads();
  unsigned a[8];
  for (int k = 0; k < 8; k++) a[k] = (pat[threadIdx.x & 63] & 16383u) * 4u;
  float acc = 0.f;
  for (int it = 0; it < iters; it += 8) {
#pragma unroll
    for (int k = 0; k < 8; k++) asm volatile("" : "+v"(a[k]));
#pragma unroll
    for (int k = 0; k < 8; k++) acc += *reinterpret_cast<lds_f *>((size_t)a[k]);
  }
  out[blockIdx.x * 256 + threadIdx.x] = acc;
}

int main() {
  std::vector<std::pair<std::string, std::vector<unsigned>>> pats;
  std::vector<unsigned> e(64);
  for (int l = 0; l < 64; l++) e[l] = l;
  pats.push_back({"consecutive", e});
  for (int l = 0; l < 64; l++) e[l] = 16 * l;
  pats.push_back({"one residue", e});
  for (int i = 0; i < 6; i++)
    for (int j = i + 1; j < 6; j++) {
      // residue = index of the coset of span{bit i, bit j} holding the lane
      std::vector<int> rep(64, -1);
      int next = 0;
      for (int l = 0; l < 64; l++) {
        const int key = l & ~((1 << i) | (1 << j));
        if (rep[key] < 0) rep[key] = next++;
        e[l] = (unsigned)rep[key] + 16u * (unsigned)l;
      }
      pats.push_back({"coset span{bit" + std::to_string(i) + ", bit" + std::to_string(j) + "}", e});
    }
  unsigned *dp = nullptr;
  float *dout = nullptr;
  if (hipMalloc(&dp, sizeof(unsigned) * 64 * pats.size()) != hipSuccess ||
      hipMalloc(&dout, sizeof(float) * 256 * 1024) != hipSuccess) {
    fprintf(stderr, "allocation failed\n");
    return 1;
  }
  for (size_t p = 0; p < pats.size(); p++)
    if (hipMemcpy(dp + 64 * p, pats[p].second.data(), sizeof(unsigned) * 64, hipMemcpyHostToDevice) != hipSuccess)
      return 1;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  const int iters = 4096, blocks = 1024;
  for (size_t p = 0; p < pats.size(); p++) {
    hipLaunchKernelGGL(k_lds_probe, dim3(blocks), dim3(256), 65536, 0, dp + 64 * p, iters, dout);
    (void)hipEventRecord(t0, 0);
    hipLaunchKernelGGL(k_lds_probe, dim3(blocks), dim3(256), 65536, 0, dp + 64 * p, iters, dout);
    (void)hipEventRecord(t1, 0);
    if (hipEventSynchronize(t1) != hipSuccess) return 1;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    // wave instructions per CU: blocks * 4 waves * iters / 256 CUs, clock 2.4 GHz
    const double cyc = ms * 1e-3 * 2.4e9 / ((double)blocks * 4 * iters / 256);
    printf("%-28s %6.2f CU-cycles per 16-byte wave read\n", pats[p].first.c_str(), cyc);
  }
  // 4-byte reads: which lanes share a conflict domain, and the modulus
  std::vector<std::pair<std::string, std::vector<unsigned>>> p32;
  for (int l = 0; l < 64; l++) e[l] = l;
  p32.push_back({"b32 consecutive", e});
  for (int l = 0; l < 64; l++) e[l] = 64 * l;
  p32.push_back({"b32 one bank (x64)", e});
  for (int l = 0; l < 64; l++) e[l] = (l % 32) + 64 * (l / 32);  // lane l, l+32 same bank, other address
  p32.push_back({"b32 halves share banks", e});
  for (int l = 0; l < 64; l++) e[l] = (l % 16) + 32 * ((l / 16) % 2) + 128 * (l / 32);  // l, l+16: equal mod 32
  p32.push_back({"b32 equal mod 32 in a half", e});
  for (int l = 0; l < 64; l++) e[l] = (l % 16) + 64 * ((l / 16) % 2) + 128 * (l / 32);  // l, l+16: equal mod 64
  p32.push_back({"b32 equal mod 64 in a half", e});
  for (int l = 0; l < 64; l++) e[l] = l + 64 * (l % 2);  // distinct banks mod 64, l and l+32 equal mod 32
  p32.push_back({"b32 distinct banks, spread", e});
  for (size_t p = 0; p < p32.size(); p++) {
    if (hipMemcpy(dp, p32[p].second.data(), sizeof(unsigned) * 64, hipMemcpyHostToDevice) != hipSuccess) return 1;
    hipLaunchKernelGGL(k_lds_probe32, dim3(blocks), dim3(256), 65536, 0, dp, iters, dout);
    (void)hipEventRecord(t0, 0);
    hipLaunchKernelGGL(k_lds_probe32, dim3(blocks), dim3(256), 65536, 0, dp, iters, dout);
    (void)hipEventRecord(t1, 0);
    if (hipEventSynchronize(t1) != hipSuccess) return 1;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    const double cyc = ms * 1e-3 * 2.4e9 / ((double)blocks * 4 * iters / 256);
    printf("%-28s %6.2f CU-cycles per 4-byte wave read\n", p32[p].first.c_str(), cyc);
  }
  return 0;
}
