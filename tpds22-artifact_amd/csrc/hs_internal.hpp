// hs_internal.hpp -- host-side internals shared by the libhs translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "hs.h"
#include "hs_device.hpp"

struct hs_arena {
  hs_arena_desc desc;
  hs::ArenaView view;      // device pointers, passed by value to kernels
  void *base;              // single device allocation holding every region
  size_t bytes;
  uint32_t *abort_host;    // hipHostMalloc'd, coherent, polled by device waits
  int device;
  int num_cus;
};

#define HS_TRY_HIP(expr)                                                           \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      fprintf(stderr, "libhs: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(_e), \
              __FILE__, __LINE__);                                                 \
      return HS_ERR_HIP;                                                           \
    }                                                                              \
  } while (0)

namespace hs_host {

// number of compute units and XCDs of the current device
int device_cus();
int device_xcds();
// groups of the XCD-hierarchical barrier for a grid of num_wgs workgroups:
// the hs_set_xcd_groups override, else the number of XCDs (doubled until no
// group has more than 255 members)
int xcd_groups(int num_wgs);
// the hs_set_xcd_groups value in force (0 = none)
int xcd_groups_override();
// co-residency check of a grid for a kernel: HS_OK or HS_ERR_RESIDENCY
int check_residency(const void *kernel, int num_wgs, int wg_size, size_t dyn_lds);
// after a persistent launch: synchronize on stream and translate the device status
int finish(hs_arena *arena, hipStream_t stream);
// clear the arena's device status word before a launch
int clear_status(hs_arena *arena, hipStream_t stream);
// initialize all barrier/lock words (fill kernels; no per-int memset)
int arena_init_words(hs_arena *arena, hipStream_t stream);
// default persistent grid: wgs_per_cu * CUs capped by residency
int default_grid(const void *kernel, int wg_size, int wgs_per_cu);
// records the current device's address of hs_scan_u32_noalloc's static scratch
int scan_scratch_init(int device);
// zero the barrier state a persistent launch of `kind` needs (capture-safe fills)
int prepare_barrier(hs_arena *ar, int kind, hipStream_t s);
// launch a persistent grid (cooperatively for HS_BAR_COOP)
int launch_persistent(const void *k, void **args, int kind, int num_wgs, int wg_size, size_t lds,
                      hipStream_t s);

}  // namespace hs_host
