"""HBM byte-counter calibration: runs the module's plain streaming copy kernel
(simk_copy_device) on a device buffer of known size, so hardware counters read
under rocprofv3 (FETCH_SIZE, WRITE_SIZE) can be scaled to bytes.

    python -m simstack_amd.calibrate --mib 1024 --reps 3
"""
import argparse
import ctypes as C
import json

from . import gpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    nbytes = args.mib << 20
    gm = gpu.GpuModule(arena_bytes=2 * nbytes + (1 << 20))
    src, dst = gm.alloc(nbytes), gm.alloc(nbytes)
    ms = C.c_float()
    times = []
    for _ in range(args.reps):
        gm._check(gm.lib.simk_copy_device(dst, src, nbytes, C.byref(ms)), "simk_copy_device")
        times.append(ms.value)
    best = min(times)
    print(json.dumps({"kernel": "simk_copy_kernel", "bytes_read": nbytes, "bytes_written": nbytes,
                      "ms": times, "GB_per_s": 2 * nbytes / (best * 1e-3) / 1e9}))


if __name__ == "__main__":
    main()
