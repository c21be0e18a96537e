"""MI355X-native NEVPT2 RDM-contraction hot path (gfx950 HIP kernels behind a C ABI)."""
