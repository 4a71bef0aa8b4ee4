"""ctypes mirrors of include/simk_types.h, simk_host.h and simk_gpu_api.h.

Struct layouts are the reference's (src/sim/kinetics/kin_mech.h:42-119,
src/gpu/gpu_api.h:30-44); test_abi.py checks the sizes against the C side.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libsimk_host.so")
GPU_LIB = os.path.join(HERE, "libsimk_gpu.so")

f64p = C.POINTER(C.c_double)


class SimkCaps(C.Structure):
    _fields_ = [("species", C.c_uint32), ("reactions", C.c_uint32), ("nnz", C.c_uint32),
                ("lu_div", C.c_uint32), ("lu_sub", C.c_uint32)]


DEFAULT_CAPS = (32, 96, 512, 512, 4096)      # kin_mech.h:22-38
# Capacities the reference's offline tools raise to (tools/ref/n2vib/n2vib_ref.cpp:17-21
# style): large enough for every synthetic mechanism used here.
RAISED_CAPS = (64, 256, 4096, 4096, 32768)


def caps(t=RAISED_CAPS):
    return SimkCaps(*t)


def mech_size(c):
    """simk_mech_layout(caps).size, restated (checked against C in test_abi)."""
    def up(x, a):
        return (x + a - 1) // a * a
    s, r, nnz, d, sub = c.species, c.reactions, c.nnz, c.lu_div, c.lu_sub
    o = 8 + 54 * r
    o = up(o, 4) + 4 + 2 * s
    o = up(o, 4) + 8 + 2 * (s + 1) + 8 * d + 4 * sub + 4 * (s + 1) + 8 * nnz
    return up(o, 4)


class KinMethod(C.Structure):
    _fields_ = [("stages", C.c_uint32), ("order", C.c_uint32), ("a", C.c_double * 6), ("c", C.c_double * 6),
                ("m", C.c_double * 4), ("e", C.c_double * 4), ("gamma0", C.c_double), ("new_f", C.c_int32 * 4),
                ("name", C.c_char_p)]


class KinControl(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("h_min", C.c_double), ("h_max", C.c_double),
                ("h_start", C.c_double), ("max_steps", C.c_uint32), ("fixed_step", C.c_int32),
                ("clamp_negative", C.c_int32)]


class GpuKinLaunch(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("y", "h", "steps", "rejects", "pulses", "failed",
                                          "k", "src", "pulse", "period", "pulse_first", "pulse_count")] + \
               [("padded", C.c_uint32), ("lane_count", C.c_uint32), ("precision", C.c_uint32),
                ("duration", C.c_double)]


class GpuInfo(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("sm_major", C.c_uint32), ("sm_minor", C.c_uint32),
                ("multiprocessors", C.c_uint32), ("total_bytes", C.c_uint64), ("free_bytes", C.c_uint64),
                ("arena_bytes", C.c_uint64)]


class SimkSynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_species", C.c_uint32), ("n_reactions", C.c_uint32),
                ("band", C.c_uint32), ("max_attempts", C.c_uint32), ("long_range", C.c_double),
                ("density_lo", C.c_double), ("density_hi", C.c_double),
                ("log10_k_lo", C.c_double), ("log10_k_hi", C.c_double), ("EaR_hi", C.c_double),
                ("T_lo", C.c_double), ("T_hi", C.c_double), ("log10_y_lo", C.c_double), ("log10_y_hi", C.c_double),
                ("pH_lo", C.c_double), ("pH_hi", C.c_double), ("acid_frac", C.c_double), ("base_frac", C.c_double),
                ("water_frac", C.c_double), ("oxidant_log10_spread", C.c_double), ("oxidant_species", C.c_uint32),
                ("pad_", C.c_uint32), ("load_log10_spread", C.c_double)]


class SimkRateLaw(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("A", C.c_double), ("n", C.c_double), ("E", C.c_double),
                ("A_inf", C.c_double), ("n_inf", C.c_double), ("fc", C.c_double), ("ref", C.c_uint32),
                ("keq_A", C.c_double), ("keq_B", C.c_double)]


class SimkAirSpecies(C.Structure):
    NAMES = ("O", "O3", "N", "NO", "NO2", "NO3", "N2O5", "N2O", "HNO2", "HNO3", "OH", "HO2", "H2O2", "H", "N2", "O2",
             "H2O")
    _fields_ = [(n, C.c_uint32) for n in NAMES]


class SimkAirSource(C.Structure):
    _fields_ = [("g_O2_diss", C.c_double), ("f_O1D", C.c_double), ("g_N2A", C.c_double), ("g_N2_diss", C.c_double),
                ("g_OH_per_100eV", C.c_double), ("boltzmann", C.c_int32), ("pad_", C.c_uint32),
                ("reduced_field_Td", C.c_double), ("filament_J_cm3", C.c_double)]


class SimkDbdZones(C.Structure):
    _fields_ = [("n_species", C.c_uint32), ("n_reactions", C.c_uint32), ("count", C.c_uint32), ("padded", C.c_uint32),
                ("block", C.c_void_p), ("T_K", C.c_void_p), ("p_Pa", C.c_void_p), ("x_H2O", C.c_void_p),
                ("power_W_cm3", C.c_void_p), ("freq_Hz", C.c_void_p), ("pulse_index", C.c_void_p)]


class SimkHybrid(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("padded", "split", "best", "steps", "cpu_samples", "gpu_samples", "dither",
                                          "pad_")] + \
               [(f, C.c_double) for f in ("cpu_lane_ms", "gpu_fixed_ms", "gpu_lane_ms", "w", "sn", "snn", "st",
                                          "snt")]


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_host = None


def host():
    """libsimk_host.so with typed entry points (raises if it is not built)."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB):
        raise RuntimeError(f"{HOST_LIB} is not built: run __graft_entry__.build()")
    lib = C.CDLL(HOST_LIB)
    V, P = C.c_void_p, C.POINTER
    _sig(lib, "simk_model_new", V)
    _sig(lib, "simk_model_free", None, V)
    _sig(lib, "simk_model_add_species", C.c_uint32, V, C.c_char_p, C.c_int32)
    _sig(lib, "simk_model_add_reaction", C.c_uint32, V, C.c_char_p, P(SimkRateLaw))
    _sig(lib, "simk_model_finalize", None, V)
    _sig(lib, "simk_model_error", C.c_int32, V, C.c_char_p, C.c_uint32)
    _sig(lib, "simk_model_species_count", C.c_uint32, V)
    _sig(lib, "simk_model_reaction_count", C.c_uint32, V)
    _sig(lib, "simk_model_mech_bytes", C.c_int32, V, P(SimkCaps), V)
    _sig(lib, "simk_model_rate_constants", None, V, C.c_double, C.c_double, f64p)
    _sig(lib, "simk_model_count_elements", None, V, f64p, f64p)
    _sig(lib, "simk_mech_build", C.c_int32, C.c_uint32, C.c_uint32, V, V, P(SimkCaps), V, C.c_char_p, C.c_uint32)
    _sig(lib, "simk_mech_stats", C.c_int32, V, P(SimkCaps), P(C.c_uint64))
    _sig(lib, "simk_method", C.c_int32, C.c_uint32, P(KinMethod))
    _sig(lib, "simk_fnv1a64", C.c_uint64, V, C.c_uint64, C.c_uint64)
    _sig(lib, "simk_zones_layout", None, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_uint64))
    _sig(lib, "simk_synth_mechanism", C.c_uint32, P(SimkSynthSpec), V, V, f64p)
    _sig(lib, "simk_synth_cells", None, P(SimkSynthSpec), f64p, C.c_uint64, C.c_uint32, C.c_uint32,
         V, V, V, V)
    _sig(lib, "simk_air_model", V, P(SimkAirSpecies))
    _sig(lib, "simk_air_source_defaults", None, P(SimkAirSource))
    _sig(lib, "simk_air_feed", C.c_double, P(SimkAirSpecies), C.c_uint32, C.c_double, C.c_double, C.c_double, f64p)
    _sig(lib, "simk_air_source_delta", C.c_int32, P(SimkAirSpecies), P(SimkAirSource), f64p, C.c_uint32, C.c_double,
         C.c_double, f64p)
    _sig(lib, "simk_dbd_fill_padding", None, P(SimkDbdZones))
    _sig(lib, "simk_dbd_reset", None, P(SimkDbdZones), P(SimkAirSpecies))
    _sig(lib, "simk_dbd_prepare_rates", None, P(SimkDbdZones), V)
    _sig(lib, "simk_dbd_prepare_source", C.c_int32, P(SimkDbdZones), P(SimkAirSpecies), P(SimkAirSource),
         C.c_uint32, C.c_double, C.c_double)
    _sig(lib, "simk_dbd_commit_pulses", None, P(SimkDbdZones))
    _sig(lib, "simk_dbd_default_control", None, P(KinControl))
    _sig(lib, "simk_hybrid_init", None, P(SimkHybrid), C.c_uint32)
    _sig(lib, "simk_hybrid_record", None, P(SimkHybrid), C.c_uint32, C.c_double, C.c_double)
    _host = lib
    return lib
