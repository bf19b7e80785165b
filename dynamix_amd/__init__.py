"""dynamix_amd -- MI355X-native XPCS g2 correlators (drop-in for dynamix's correlators).

The correlators live in dynamix_amd.correlator (dense, event, fft); the HIP kernels and
their C ABI are built into dynamix_amd/libdxg2.so by dynamix_amd.build.
"""
version = "0.1.0"
