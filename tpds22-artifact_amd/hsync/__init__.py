"""hsync -- MI355X device-wide synchronization primitives and persistent kernels.

Python front end of libhs.so (C ABI in include/hs.h, HIP kernels for gfx950).
"""
from ._lib import (BAR_COOP, BAR_FLAT, BAR_LF, BAR_NAIVE_ALL, BAR_TREE, BAR_TWO_LEVEL, BAR_XCD,
                   BAR_XCD_FLAT_RELEASE, BARRIER_NAMES, HSError, MUTEX_EBO, MUTEX_FA, MUTEX_SLEEP,
                   MUTEX_SPIN, SEM_EBO, SEM_EBO_PRIORITY, SEM_SPIN, SEM_SPIN_PRIORITY, SEM_XCD_QUEUE,
                   LIB_PATH, load)
from .api import (PRIMS, Arena, barrier_bench, barrier_kind, check_words, device_info,
                  gen_uniform_u32, gen_uniform_u32_host, max_resident_wgs, num_locs_mult, prim_id,
                  probe_xcc_ids, reduce, reduce_batch_i32, reduce_i32, reduce_i32_stamped, stamp_breakdown, run_sync_prim, scan_batch_u32,
                  scan_u32, scan_u32_noalloc, sem_bench, sem_wait_stats,
                  get_scan_round_bytes, set_scan_round_bytes, set_xcd_groups, storage_locs,
                  sum_scan_blelloch, sync_bench, tune_barrier, tune_reduce_i32, tune_scan_batch_u32, tune_scan_u32)

__all__ = [n for n in dir() if not n.startswith("_")]
