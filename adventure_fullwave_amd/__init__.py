"""MI355X-native HDDM full-wave hot path (ADVENTURE_FullWave drop-in).

The product is the HIP library libfwhip.so (C ABI in include/fwhip.h);
this package holds its ctypes front end and the synthetic model generator.
"""
