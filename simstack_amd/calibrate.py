"""HBM byte-counter calibration: runs the module's plain streaming copy kernel
on a device buffer of known size, so hardware counters read under rocprofv3
(FETCH_SIZE, WRITE_SIZE) can be scaled to bytes. --width 16 copies with one
16-byte access per lane (simk_copy_device), --width 8 with one 8-byte access
per lane (simk_copy_device_b64, the access width of the solver and advection
kernels; the counters scale differently for the two widths).

    python -m simstack_amd.calibrate --mib 1024 --reps 3 [--width 8]
"""
import argparse
import ctypes as C
import json

from . import gpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--width", type=int, default=16, choices=(8, 16))
    args = ap.parse_args()
    nbytes = args.mib << 20
    gm = gpu.GpuModule(arena_bytes=2 * nbytes + (1 << 20))
    src, dst = gm.alloc(nbytes), gm.alloc(nbytes)
    ms = C.c_float()
    times = []
    for _ in range(args.reps):
        fn = gm.lib.simk_copy_device if args.width == 16 else gm.lib.simk_copy_device_b64
        gm._check(fn(dst, src, nbytes, C.byref(ms)), "copy")
        times.append(ms.value)
    best = min(times)
    print(json.dumps({"kernel": "simk_copy_kernel" if args.width == 16 else "simk_copy8_kernel",
                      "bytes_read": nbytes, "bytes_written": nbytes,
                      "ms": times, "GB_per_s": 2 * nbytes / (best * 1e-3) / 1e9}))


if __name__ == "__main__":
    main()
