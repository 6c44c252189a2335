"""glide_amd - MI355X-native SSA / FASCD ice-flow solver (HIP kernels behind a C ABI)."""
