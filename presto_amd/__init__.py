"""MI355X-native accelsearch hot path: HIP kernels behind a C ABI
(include/presto_accel.h), with a ctypes binding for Python callers."""
