// Launch-shape defaults of libfocus_gpu.so on MI355X (gfx950).  Rewritten by
// `python focus-stacker_amd/autotune.py --write` from its measurements on the GPU; the figures it
// measured are kept beside it in focus-stacker_amd/tuning_mi355x.json.
#pragma once

// Single stacks: a pyramid level runs 64x16 tiles while it has at least this many of them, then
// 16x16 tiles, then (fewer 16x16 tiles than this) one frame per block with an ordered sum.
#define FG_TUNE_PASS2_MIN_TILES 512
// Batches: a level's launch uses 64x16 tiles when the stacks' mean width at that level is at least
// this, else 16x16 tiles.
#define FG_TUNE_BATCH_WIDE_LEVEL 512
// Batches: persistent pass-1 blocks per CU (0: as many as fit).
#define FG_TUNE_BATCH_PASS1_BLOCKS_PER_CU 0
// Batches: a group holds at most this many MiB of stack buffers (0: the memory budget); smaller
// groups keep each stack's pass 2 close in time to its pass 1.
#define FG_TUNE_BATCH_GROUP_MIB 0
// Batches: streams of a batch: the groups run in order on the first, stacks of their own alternate
// over the others (their launches interleave with each other and with the groups).
#define FG_TUNE_BATCH_STREAMS 3
// Batches: a pyramid stack whose output has at least this many thousand pixels runs as a stack of its
// own (ordinary launches on the other streams, beside the groups), since it fills the GPU alone
// (0: every pyramid stack joins a group).
#define FG_TUNE_BATCH_LARGE_KPIX 4000
// Single stacks: each XCD runs a contiguous run of pass-1 / pass-2 tiles, ordered in strips of
// this many tile rows (compile time; 1: row-major).  Measured on the 143 x 7952x5304 u16 stack:
// row-major runs 117.0 ms, blocks dealt round-robin 120.7 ms, strips of 4-16 rows 119.9-121.2 ms
// (taller strips slow pass 1).
#ifndef FG_TUNE_XCD_STRIP
#define FG_TUNE_XCD_STRIP 1
#endif
