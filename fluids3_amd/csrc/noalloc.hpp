// noalloc.hpp -- the scratch array behind fx_scan_device_noalloc.
//
// The scan's two upper levels (block totals and their scans) live in a
// file-scope device array of noalloc.hip instead of device allocations:
// FX_SCAN_L2 words for level-2 totals and their scan, FX_SCAN_L3 for level 3.
#pragma once

#include "fluidsx.h"

#define FX_SCAN_L2 (FX_SCAN_NOALLOC_MAX / 1024 + 1)
#define FX_SCAN_L3 (FX_SCAN_L2 / 1024 + 1)
#define FX_SCAN_SCRATCH_WORDS (2 * FX_SCAN_L2 + 2 * FX_SCAN_L3)
#define FX_MAX_DEVICES 64

namespace fx {
// Device address of the current device's scratch array (resolved once per
// device with hipGetSymbolAddress, then cached).
int scratch_setup(void** out);
}  // namespace fx
