"""MI355X-native LocARNA sequence-structure alignment (gfx950 HIP kernels
behind a C ABI; see include/locarna_amd.h)."""
