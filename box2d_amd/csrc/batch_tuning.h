// Configuration of the batched persistent solver per world-size bucket, chosen offline on an
// MI355X by box2d_amd/autotune.py (which rewrites this file): for each bucket, the workgroup size
// and warm-start form with the lowest median GPU time of the substep-loop launch over a batch of
// worlds of that size. A batch's launch takes the size of the bucket with the most worlds; each
// world its bucket's warm-start form. The results do not depend on these values, only the time.
#pragma once

// { max awake bodies, workgroup threads, warm start body by body }
#define B2AMD_BATCH_TUNING_TABLE                                                                                         \
	{ { 64, 256, 1 }, { 256, 128, 0 }, { 1024, 512, 0 }, { 5400, 1024, 0 } }
#define B2AMD_BATCH_TUNING_SOURCE "box2d_amd/autotune.py, 2026-12-03"
