"""The GPU module (libsimk_gpu.so) from Python: the reference's GPU API
(src/gpu/gpu_api.h) plus the extensions in include/simk_gpu_api.h.

There is deliberately no CPU fallback: GpuModule() raises if the library is
not built or no device is present, and every solver call goes to the device.
"""
import ctypes as C
import os

import numpy as np

from . import _ffi

V = C.c_void_p


class SimkGenOptions(C.Structure):
    _fields_ = [("fence_rows", C.c_uint32), ("waves_per_eu", C.c_uint32), ("unroll_stages", C.c_uint32),
                ("reg_window", C.c_uint32), ("batch", C.c_uint32), ("precision", C.c_uint32),
                ("sweeps", C.c_uint32), ("reuse", C.c_uint32), ("lds_slots", C.c_uint32),
                ("fun_in_jac", C.c_uint32), ("pair_slots", C.c_uint32), ("k_pairs", C.c_uint32),
                ("stage_carry", C.c_uint32)]


class SimkKinInfo(C.Structure):
    _fields_ = [("kernel", C.c_uint32), ("ws_slots", C.c_uint32), ("waves", C.c_uint32), ("ordered", C.c_uint32),
                ("species", C.c_uint32),
                ("reactions", C.c_uint32), ("nnz", C.c_uint32), ("lu_div", C.c_uint32), ("lu_sub", C.c_uint32),
                ("jit_cached", C.c_uint32), ("jit_ms", C.c_double), ("last_kernel_ms", C.c_double),
                ("signature", C.c_uint64), ("kernel_name", C.c_char * 64)]


class SimkTuneResult(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("lanes_per_cell", C.c_uint32), ("waves", C.c_uint32), ("batch", C.c_uint32),
                ("reg_window", C.c_uint32), ("unroll_stages", C.c_uint32), ("reuse", C.c_uint32),
                ("lds_slots", C.c_uint32), ("candidates", C.c_uint32),
                ("sample_cells", C.c_uint32), ("best_ms", C.c_float), ("default_ms", C.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SimkServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_requests", C.c_uint32), ("steps", C.c_uint32),
                ("substeps", C.c_uint32), ("slice_us", C.c_uint32), ("dt", C.c_double)]


class SimkServeRequest(C.Structure):
    _fields_ = [("mech", C.c_uint32), ("cells", C.c_uint32), ("channels", C.c_uint32), ("nodes", C.c_uint32),
                ("padded", C.c_uint32), ("kind", C.c_uint32), ("block", C.c_void_p), ("courant", C.c_void_p),
                ("inflow", C.c_void_p), ("result", C.c_void_p), ("steps", C.c_uint32), ("precision", C.c_uint32),
                ("dbd", C.c_void_p)]


class SimkServeDbd(C.Structure):
    _fields_ = [("species", _ffi.SimkAirSpecies), ("mode", C.c_uint32), ("source", _ffi.SimkAirSource),
                ("residence_s", C.c_double), ("T_K", C.c_void_p), ("power_W_cm3", C.c_void_p),
                ("freq_Hz", C.c_void_p), ("pulse_index", C.c_void_p)]


class SimkServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("queued", C.c_uint32), ("active", C.c_uint32), ("iterations", C.c_uint64),
                ("admitted_bytes", C.c_uint64), ("peak_admitted_bytes", C.c_uint64), ("advect_ms", C.c_double),
                ("chem_ms", C.c_double), ("copy_ms", C.c_double), ("carried", C.c_uint64)]


# Every export the reference's loader resolves (src/gpu/gpu_api.h:112-115).
REFERENCE_EXPORTS = ("gpu_init", "gpu_error_string", "gpu_alloc", "gpu_arena_used", "gpu_arena_reset",
                     "gpu_copy_to_device", "gpu_copy_to_host", "gpu_sync", "gpu_step_points", "gpu_kin_set_mech",
                     "gpu_kin_advance", "gpu_host_register", "gpu_host_unregister", "gpu_kin_advance_async",
                     "gpu_kin_wait", "gpu_kin_set_mode")
EXTENSION_EXPORTS = ("simk_kin_set_mech_caps", "simk_kin_advance_n", "simk_kin_advance_static", "simk_kin_set_team", "simk_kin_set_grid", "simk_kin_set_order",
                     "simk_kin_set_queue_policy",
                     "simk_kin_set_kernel", "simk_kin_set_gen_options", "simk_kin_info",
                     "simk_jit_precompile", "simk_copy_device", "simk_copy_device_b64", "simk_upwind_f64",
                     "simk_upwind_f32",
                     "simk_kin_set_mechs", "simk_kin_advance_multi", "simk_graph_capture_multi", "simk_graph_replay",
                     "simk_graph_destroy", "simk_jit_precompile_multi", "simk_serve_open", "simk_serve_submit",
                     "simk_serve_step", "simk_serve_stats", "simk_serve_close", "simk_kin_set_cache_budget",
                     "simk_kin_autotune", "simk_kin_autotune_multi", "simk_plugflow_f64",
                     "simk_kin_set_mech_control")


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


def load_library(path=_ffi.GPU_LIB):
    """dlopen + typed symbols, without touching a device (safe on a CPU host)."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.POINTER
    _sig(lib, "gpu_init", C.c_int32, C.c_uint32, C.c_uint64, P(_ffi.GpuInfo))
    _sig(lib, "gpu_error_string", C.c_char_p, C.c_int32)
    _sig(lib, "gpu_alloc", V, C.c_uint64)
    _sig(lib, "gpu_arena_used", C.c_uint64)
    _sig(lib, "gpu_arena_reset", None)
    _sig(lib, "gpu_copy_to_device", C.c_int32, V, V, C.c_uint64)
    _sig(lib, "gpu_copy_to_host", C.c_int32, V, V, C.c_uint64)
    _sig(lib, "gpu_sync", C.c_int32)
    _sig(lib, "gpu_step_points", C.c_int32, V, V, V, V, V, V, C.c_uint32, C.c_float)
    _sig(lib, "gpu_kin_set_mech", C.c_int32, V, P(_ffi.KinMethod), P(_ffi.KinControl))
    _sig(lib, "gpu_kin_advance", C.c_int32, P(_ffi.GpuKinLaunch), P(C.c_float))
    _sig(lib, "gpu_host_register", C.c_int32, V, C.c_uint64)
    _sig(lib, "gpu_host_unregister", C.c_int32, V)
    _sig(lib, "gpu_kin_advance_async", C.c_int32, P(_ffi.GpuKinLaunch), V, V, C.c_uint64, C.c_uint64)
    _sig(lib, "gpu_kin_wait", C.c_int32, P(C.c_float), P(C.c_float))
    _sig(lib, "gpu_kin_set_mode", C.c_int32, C.c_uint32)
    _sig(lib, "simk_kin_set_mech_caps", C.c_int32, V, P(_ffi.SimkCaps), P(_ffi.KinMethod), P(_ffi.KinControl))
    _sig(lib, "simk_kin_set_kernel", C.c_int32, C.c_uint32)
    _sig(lib, "simk_kin_set_gen_options", C.c_int32, P(SimkGenOptions))
    _sig(lib, "simk_kin_info", C.c_int32, P(SimkKinInfo))
    _sig(lib, "simk_jit_precompile", C.c_int32, V, P(_ffi.SimkCaps), P(_ffi.KinMethod), P(SimkGenOptions),
         C.c_char_p, C.c_uint32)
    _sig(lib, "simk_copy_device", C.c_int32, V, V, C.c_uint64, P(C.c_float))
    _sig(lib, "simk_copy_device_b64", C.c_int32, V, V, C.c_uint64, P(C.c_float))
    _sig(lib, "simk_kin_advance_n", C.c_int32, P(_ffi.GpuKinLaunch), C.c_uint32, P(C.c_float))
    _sig(lib, "simk_kin_advance_static", C.c_int32, P(_ffi.GpuKinLaunch), C.c_uint32, P(C.c_float))
    _sig(lib, "simk_kin_set_grid", C.c_int32, C.c_uint32)
    _sig(lib, "simk_kin_set_team", C.c_int32, C.c_uint32)
    _sig(lib, "simk_kin_set_order", C.c_int32, V)
    _sig(lib, "simk_kin_set_queue_policy", C.c_int32, C.c_uint32)
    _sig(lib, "simk_kin_set_mechs", C.c_int32, C.c_uint32, V, V, P(_ffi.KinMethod), P(_ffi.KinControl))
    _sig(lib, "simk_jit_precompile_multi", C.c_int32, C.c_uint32, V, V, P(_ffi.KinMethod), P(SimkGenOptions),
         C.c_char_p, C.c_uint32)
    _sig(lib, "simk_kin_advance_multi", C.c_int32, C.c_uint32, V, C.c_uint32, P(C.c_float))
    _sig(lib, "simk_graph_capture_multi", C.c_int32, C.c_uint32, V, C.c_uint32, C.c_uint32, P(C.c_uint32))
    _sig(lib, "simk_graph_replay", C.c_int32, C.c_uint32, P(C.c_float))
    _sig(lib, "simk_graph_destroy", C.c_int32, C.c_uint32)
    _sig(lib, "simk_upwind_f64", C.c_int32, V, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, V, C.c_uint32, V)
    _sig(lib, "simk_upwind_f32", C.c_int32, V, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, V, C.c_uint32, V)
    _sig(lib, "simk_plugflow_f64", C.c_int32, V, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, V)
    _sig(lib, "simk_kin_set_cache_budget", C.c_int32, C.c_uint64)
    _sig(lib, "simk_kin_autotune", C.c_int32, P(_ffi.GpuKinLaunch), C.c_uint32, C.c_double, P(SimkTuneResult))
    _sig(lib, "simk_kin_autotune_multi", C.c_int32, C.c_uint32, P(_ffi.GpuKinLaunch), C.c_uint32, C.c_double,
         P(SimkTuneResult))
    _sig(lib, "simk_serve_open", C.c_int32, P(SimkServeConfig))
    _sig(lib, "simk_serve_submit", C.c_int32, P(SimkServeRequest), P(C.c_uint32))
    _sig(lib, "simk_kin_set_mech_control", C.c_int32, C.c_uint32, P(_ffi.KinControl))
    _sig(lib, "simk_serve_step", C.c_int32, P(C.c_uint32), C.c_uint32, P(C.c_uint32))
    _sig(lib, "simk_serve_stats", C.c_int32, P(SimkServeStats))
    _sig(lib, "simk_serve_close", C.c_int32)
    return lib


def precompile(mech, meth, options=None, lib=None):
    """Generates and compiles a mechanism's kernel into the on-disk cache (no device needed)."""
    lib = lib or load_library()
    log = C.create_string_buffer(1 << 16)
    opt = options or SimkGenOptions()
    rc = lib.simk_jit_precompile(mech.bytes.ctypes.data, C.byref(mech.caps), C.byref(meth), C.byref(opt), log,
                                 1 << 16)
    if rc != 0:
        raise RuntimeError(f"kernel compile failed ({rc}):\n{log.value.decode(errors='replace')[-4000:]}")
    return log.value.decode(errors="replace")


def precompile_multi(mechs, meth, options=None, lib=None):
    """precompile() for the multi-mechanism kernel of GpuModule.set_mechs."""
    lib = lib or load_library()
    log = C.create_string_buffer(1 << 16)
    opt = options or SimkGenOptions()
    ptrs = (C.c_void_p * len(mechs))(*[m.bytes.ctypes.data for m in mechs])
    caps = (_ffi.SimkCaps * len(mechs))(*[m.caps for m in mechs])
    rc = lib.simk_jit_precompile_multi(len(mechs), ptrs, caps, C.byref(meth), C.byref(opt), log, 1 << 16)
    if rc != 0:
        raise RuntimeError(f"kernel compile failed ({rc}):\n{log.value.decode(errors='replace')[-4000:]}")
    return log.value.decode(errors="replace")


class GpuError(RuntimeError):
    pass


class GpuModule:
    """One initialised GPU module: arena, stream, current mechanism."""

    def __init__(self, arena_bytes=1 << 30, device=0, path=_ffi.GPU_LIB):
        self.lib = load_library(path)
        self.info = _ffi.GpuInfo()
        self._check(self.lib.gpu_init(device, arena_bytes, C.byref(self.info)), "gpu_init")

    def _check(self, rc, what):
        if rc != 0:
            raise GpuError(f"{what}: {self.lib.gpu_error_string(rc).decode()} ({rc})")

    def set_mech(self, mech, meth, ctl):
        self._check(self.lib.simk_kin_set_mech_caps(mech.bytes.ctypes.data, C.byref(mech.caps), C.byref(meth),
                                                    C.byref(ctl)), "simk_kin_set_mech_caps")

    def set_kernel(self, kind):
        self._check(self.lib.simk_kin_set_kernel(kind), "simk_kin_set_kernel")

    def set_gen_options(self, fence_rows=0, waves_per_eu=0, unroll_stages=0, reg_window=0, batch=0, sweeps=0,
                        reuse=0, lds_slots=0, fun_in_jac=0, pair_slots=0, k_pairs=0, stage_carry=0):
        opt = SimkGenOptions(fence_rows, waves_per_eu, unroll_stages, reg_window, batch, 0, sweeps, reuse, lds_slots,
                             fun_in_jac, pair_slots, k_pairs, stage_carry)
        self._check(self.lib.simk_kin_set_gen_options(C.byref(opt)), "simk_kin_set_gen_options")

    def set_mechs(self, mechs, meth, ctl):
        """Ragged batch: up to four mechanisms compiled into one persistent kernel."""
        ptrs = (C.c_void_p * len(mechs))(*[m.bytes.ctypes.data for m in mechs])
        caps = (_ffi.SimkCaps * len(mechs))(*[m.caps for m in mechs])
        self._check(self.lib.simk_kin_set_mechs(len(mechs), ptrs, caps, C.byref(meth), C.byref(ctl)),
                    "simk_kin_set_mechs")

    def advance_multi(self, dzs, duration, repeats=1, precision=0):
        """One launch advancing every DeviceZones in dzs (mechanism i = dzs[i]);
        precision 0 (FP64) or 3 (FP32 factor storage)."""
        Ls = (_ffi.GpuKinLaunch * len(dzs))(*[d.launch_desc(duration, precision=precision) for d in dzs])
        ms = C.c_float()
        self._check(self.lib.simk_kin_advance_multi(len(dzs), Ls, repeats, C.byref(ms)), "simk_kin_advance_multi")
        return ms.value

    def autotune_multi(self, dzs, duration=0.0, sample_cells=0):
        """simk_kin_autotune_multi over the blocks dzs (mechanism i = dzs[i])."""
        Ls = (_ffi.GpuKinLaunch * len(dzs))(*[d.launch_desc(duration) for d in dzs])
        res = SimkTuneResult()
        self._check(self.lib.simk_kin_autotune_multi(len(dzs), Ls, sample_cells, duration, C.byref(res)),
                    "simk_kin_autotune_multi")
        return res

    def capture_multi(self, dzs, duration, repeats, steps):
        Ls = (_ffi.GpuKinLaunch * len(dzs))(*[d.launch_desc(duration) for d in dzs])
        gid = C.c_uint32()
        self._check(self.lib.simk_graph_capture_multi(len(dzs), Ls, repeats, steps, C.byref(gid)), "capture")
        return gid.value

    def destroy_graph(self, gid):
        self._check(self.lib.simk_graph_destroy(gid), "simk_graph_destroy")

    def replay(self, gid):
        ms = C.c_float()
        self._check(self.lib.simk_graph_replay(gid, C.byref(ms)), "simk_graph_replay")
        return ms.value

    # ---- serving loop (simk_serve_*) ----
    def serve_open(self, hbm_budget=0, max_requests=0, steps=0, substeps=0, dt=0.0, slice_us=0):
        cfg = SimkServeConfig(hbm_budget, max_requests, steps, substeps, slice_us, dt)
        self._check(self.lib.simk_serve_open(C.byref(cfg)), "simk_serve_open")
        self._serve_keep = {}

    def serve_submit(self, mech_index, block_ptr, padded, cells, channels, nodes, courant, result, inflow=None,
                     keep=None, precision=0):
        """Queues a request whose input zone block starts at host address
        block_ptr (row stride padded); result: float64 [species + 5, cells].
        precision 1: f32 delta lanes."""
        courant = np.ascontiguousarray(courant, dtype=np.float64)
        inflow = None if inflow is None else np.ascontiguousarray(inflow, dtype=np.float64)
        assert result.dtype == np.float64 and result.flags.c_contiguous and result.shape[1] == cells
        req = SimkServeRequest(mech_index, cells, channels, nodes, padded, 0, block_ptr, courant.ctypes.data,
                               None if inflow is None else inflow.ctypes.data, result.ctypes.data, 0, precision, None)
        t = C.c_uint32()
        self._check(self.lib.simk_serve_submit(C.byref(req), C.byref(t)), "simk_serve_submit")
        self._serve_keep[t.value] = (courant, inflow, result, keep)
        return t.value

    def serve_submit_dbd(self, mech_index, zones, steps, residence=2.0, mode=0, source=None, feed=None, species=None,
                         precision=0):
        """Queues a DBD reactor request: a dbd.DbdZones block (its operating
        point and pulse clock), `steps` master steps, plug flow over
        `residence` seconds toward `feed` (0: independent zones). `species`:
        the air species map (default: the block's mechanism's); precision 1:
        f32 delta lanes (the reference scene's default). Returns (ticket,
        result array [species + 5, zones])."""
        from . import dbd as _dbd
        result = np.zeros((zones.n + 5, zones.count))
        arrays = [np.ascontiguousarray(getattr(zones, a)[:zones.count], dtype=np.float64)
                  for a in ("T_K", "power_W_cm3", "freq_Hz", "pulse_index")]
        d = SimkServeDbd(species if species is not None else zones.mech.sp, mode, source if source is not None else _dbd.source_defaults(), residence,
                         *[a.ctypes.data for a in arrays])
        feed = None if feed is None else np.ascontiguousarray(feed, dtype=np.float64)
        req = SimkServeRequest(mech_index, zones.count, 1, zones.count, zones.padded, 1, zones.block.ctypes.data, None,
                               None if feed is None else feed.ctypes.data, result.ctypes.data, steps, precision,
                               C.addressof(d))
        t = C.c_uint32()
        self._check(self.lib.simk_serve_submit(C.byref(req), C.byref(t)), "simk_serve_submit")
        self._serve_keep[t.value] = (feed, result, zones)
        return t.value, result

    def set_mech_control(self, index, ctl):
        self._check(self.lib.simk_kin_set_mech_control(index, C.byref(ctl)), "simk_kin_set_mech_control")

    def serve_submit_zones(self, mech_index, zones, channels, nodes, courant, inflow=None, precision=0):
        """serve_submit of a whole problem.Zones block; returns (ticket, result array)."""
        result = np.zeros((zones.n + 5, zones.count))
        t = self.serve_submit(mech_index, zones.block.ctypes.data, zones.padded, zones.count, channels, nodes,
                              courant, result, inflow, keep=zones, precision=precision)
        return t, result

    def serve_step(self, max_done=4096):
        done = (C.c_uint32 * max_done)()
        n = C.c_uint32()
        self._check(self.lib.simk_serve_step(done, max_done, C.byref(n)), "simk_serve_step")
        out = [done[i] for i in range(n.value)]
        for t in out:
            self._serve_keep.pop(t, None)
        return out

    def serve_stats(self):
        st = SimkServeStats()
        self._check(self.lib.simk_serve_stats(C.byref(st)), "simk_serve_stats")
        return st

    def serve_close(self):
        self._check(self.lib.simk_serve_close(), "simk_serve_close")

    def set_cache_budget(self, nbytes):
        self._check(self.lib.simk_kin_set_cache_budget(nbytes), "simk_kin_set_cache_budget")

    def set_grid(self, waves):
        self._check(self.lib.simk_kin_set_grid(waves), "simk_kin_set_grid")

    def set_queue_policy(self, policy):
        """1 (default): advances of the same block take cells longest first by
        the previous advance's attempts; 0: natural order."""
        self._check(self.lib.simk_kin_set_queue_policy(policy), "simk_kin_set_queue_policy")

    def kin_info(self):
        info = SimkKinInfo()
        self._check(self.lib.simk_kin_info(C.byref(info)), "simk_kin_info")
        return info

    def alloc(self, nbytes):
        p = self.lib.gpu_alloc(nbytes)
        if not p:
            raise GpuError(f"gpu_alloc({nbytes}): arena full")
        return p

    def arena_reset(self):
        self.lib.gpu_arena_reset()

    def to_device(self, dev, host_array):
        a = np.ascontiguousarray(host_array)
        self._check(self.lib.gpu_copy_to_device(dev, a.ctypes.data, a.nbytes), "gpu_copy_to_device")

    def to_host(self, host_array, dev, nbytes=None):
        self._check(self.lib.gpu_copy_to_host(host_array.ctypes.data, dev, nbytes or host_array.nbytes),
                    "gpu_copy_to_host")


class DeviceZones:
    """A zone block mirrored in the module's arena (dbd_advance_gpu's DbdGpu,
    src/sim/dbd.cpp:263-295): device pointer = device base + host row offset."""

    def __init__(self, gm, zones):
        self.gm, self.z = gm, zones
        self.d_block = gm.alloc(zones.block.nbytes)

    def upload(self):
        self.gm.to_device(self.d_block, self.z.block)

    def download(self, mutable_only=True):
        n = self.z.mutable * 8 if mutable_only else self.z.block.nbytes
        self.gm.to_host(self.z.block, self.d_block, n)

    def launch_desc(self, duration, lane_count=0, precision=0):
        L = _ffi.GpuKinLaunch()
        for name, off in zip(("y", "h", "steps", "rejects", "pulses", "failed", "k", "src", "pulse", "period",
                              "pulse_first", "pulse_count"), self.z.offsets):
            setattr(L, name, self.d_block + off * 8)
        L.padded = self.z.padded
        L.lane_count = lane_count or self.z.count
        L.precision = precision
        L.duration = duration
        return L

    def autotune(self, duration=0.0, sample_cells=0, precision=0):
        """simk_kin_autotune on this block (a device copy of it); the chosen
        settings stay in effect for later advances of that precision (0 FP64,
        2 mixed)."""
        L = self.launch_desc(duration or 1.0, precision=precision)
        res = SimkTuneResult()
        self.gm._check(self.gm.lib.simk_kin_autotune(C.byref(L), sample_cells, duration, C.byref(res)),
                       "simk_kin_autotune")
        return res

    def advance(self, duration, lane_count=0, precision=0, repeats=1):
        """gpu_kin_advance (or, repeats > 1, simk_kin_advance_n) on the device block;
        returns kernel milliseconds."""
        L = self.launch_desc(duration, lane_count, precision)
        ms = C.c_float(0.0)
        if repeats == 1:
            self.gm._check(self.gm.lib.gpu_kin_advance(C.byref(L), C.byref(ms)), "gpu_kin_advance")
        else:
            self.gm._check(self.gm.lib.simk_kin_advance_n(C.byref(L), repeats, C.byref(ms)), "simk_kin_advance_n")
        return ms.value
