// scan -- scan/main.cu on libhs: ./scan N [inclusive]
// All-ones input, CPU reference scan (timed), GPU scan, "Match: 1/0" with a
// diff window on mismatch (scan/main.cu:23-114); exits non-zero on mismatch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iostream>
#include <vector>

#include "hip_check.hpp"
#include "hs.h"

static void cpu_sum_scan(uint32_t *out, const uint32_t *in, size_t n, bool inclusive) {
  uint32_t run = 0;
  for (size_t i = 0; i < n; ++i) {
    if (inclusive) { run += in[i]; out[i] = run; }
    else { out[i] = run; run += in[i]; }
  }
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s N [inclusive]\n", argv[0]);
    return 2;
  }
  const size_t n = strtoull(argv[1], nullptr, 10);
  const bool inclusive = argc > 2 && atoi(argv[2]) != 0;
  std::cout << "h_in size: " << n << std::endl;
  std::vector<uint32_t> in(n, 1), ref(n), got(n);
  std::clock_t t0 = std::clock();
  cpu_sum_scan(ref.data(), in.data(), n, inclusive);
  std::cout << "CPU time: " << (std::clock() - t0) / (double)CLOCKS_PER_SEC << std::endl;
  hs_device_info info;
  if (hs_device_info_get(&info)) return 1;
  hs_arena_desc d = {8 * info.num_cus, 1, 1, info.num_cus, 0, 0};
  hs_arena *a = nullptr;
  if (hs_arena_create(&d, &a)) return 1;
  uint32_t *din = nullptr, *dout = nullptr, *scr = nullptr;
  uint64_t ns = 0;
  hs_scan_scratch_size(0, n, &ns);
  if (hipMalloc(&din, 4 * (n ? n : 1)) || hipMalloc(&dout, 4 * (n ? n : 1)) || hipMalloc(&scr, 4 * ns))
    return 1;
  HIP_CHECK(hipMemcpy(din, in.data(), 4 * n, hipMemcpyHostToDevice));
  int rc = hs_scan_u32(a, din, dout, n, inclusive, scr, HS_BAR_XCD, 0, nullptr);
  if (!rc) rc = hs_arena_sync(a, nullptr);
  if (rc) {
    fprintf(stderr, "ERROR: %s\n", hs_status_string(rc));
    return 1;
  }
  HIP_CHECK(hipMemcpy(got.data(), dout, 4 * n, hipMemcpyDeviceToHost));
  bool match = true;
  size_t diff = 0;
  for (size_t i = 0; i < n; ++i)
    if (ref[i] != got[i]) { match = false; diff = i; break; }
  std::cout << "Match: " << match << std::endl;
  if (!match) {
    std::cout << "Difference in index: " << diff << std::endl;
    std::cout << "CPU: " << ref[diff] << std::endl << "GPU: " << got[diff] << std::endl;
  }
  HIP_CHECK(hipFree(din));
  HIP_CHECK(hipFree(dout));
  HIP_CHECK(hipFree(scr));
  hs_arena_destroy(a);
  return match ? 0 : 1;
}
