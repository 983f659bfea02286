"""MI355X-native LDU solve path (HIP kernels behind a C-ABI; see include/unap_hip.h)."""
