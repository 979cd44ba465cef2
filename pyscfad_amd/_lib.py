"""ctypes binding of libpyscfad_amd.so (include/pyscfad_amd.h).

Loading fails loudly: there is no CPU fallback for the product path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyscfad_amd.so")

PREC_F64 = 0
PREC_F32 = 1
PREC_BF16X3 = 2
PREC_BF16X6 = 3
PRECISIONS = {"f64": PREC_F64, "f32": PREC_F32, "bf16x3": PREC_BF16X3, "bf16x6": PREC_BF16X6}

# every symbol include/pyscfad_amd.h declares
EXPORTS = (
    "pyscfad_amd_ccsd_t_energy",
    "pyscfad_amd_ccsd_t_energy_vjp",
    "pyscfad_amd_ccsd_t_energy_vjp_noalloc",
    "pyscfad_amd_ccsd_t_workspace_bytes",
    "pyscfad_amd_ccsd_t_num_triples",
    "pyscfad_amd_ccsd_t_energy_slice",
    "pyscfad_amd_ccsd_t_energy_vjp_slice",
    "pyscfad_amd_ccsd_t_batched_workspace_bytes",
    "pyscfad_amd_ccsd_t_batched",
    "pyscfad_amd_ccsd_t_autotune",
    "pyscfad_amd_ccsd_t_df_workspace_bytes",
    "pyscfad_amd_ccsd_t_df_energy",
    "pyscfad_amd_ccsd_t_df_energy_vjp",
    "pyscfad_amd_lno_ccsd_t_workspace_bytes",
    "pyscfad_amd_lno_ccsd_t_energy",
    "pyscfad_amd_lno_ccsd_t_energy_vjp",
    "pyscfad_amd_last_error",
    "pyscfad_amd_set_timing",
    "pyscfad_amd_last_timing",
    "pyscfad_amd_xla_ccsd_t_energy",
    "pyscfad_amd_xla_ccsd_t_energy_vjp",
    "pyscfad_amd_graph_cache_stats",
    "pyscfad_amd_abi_version",
)

_lib = None


class Instance(ctypes.Structure):
    """PyscfadAmdInstance of include/pyscfad_amd.h (one problem of a batched call)."""
    _fields_ = ([("nocc", ctypes.c_int), ("nvir", ctypes.c_int), ("ovvv_packed", ctypes.c_int),
                 ("et_bar", ctypes.c_double)]
                + [(n, ctypes.c_void_p) for n in ("mo_energy", "t1", "t2", "fock", "ovoo", "ovov", "ovvv",
                                                  "t1_bar", "t2_bar", "fock_bar", "mo_energy_bar", "ovoo_bar",
                                                  "ovov_bar", "ovvv_bar", "e_out")])


class T3Error(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise T3Error(f"{LIB_PATH} is missing: build it with `python -m pyscfad_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, dp, i, d, sz = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    lib.pyscfad_amd_ccsd_t_energy.restype = i
    lib.pyscfad_amd_ccsd_t_energy.argtypes = [i, i] + [dp] * 7 + [i, i, dp, vp, sz, vp]
    lib.pyscfad_amd_ccsd_t_energy_vjp.restype = i
    lib.pyscfad_amd_ccsd_t_energy_vjp.argtypes = [i, i] + [dp] * 7 + [i, d, i] + [dp] * 8 + [vp, sz, vp]
    lib.pyscfad_amd_ccsd_t_energy_vjp_noalloc.restype = i
    lib.pyscfad_amd_ccsd_t_energy_vjp_noalloc.argtypes = [i, i] + [dp] * 7 + [i, d, i] + [dp] * 8 + [vp]
    lib.pyscfad_amd_ccsd_t_workspace_bytes.restype = sz
    lib.pyscfad_amd_ccsd_t_workspace_bytes.argtypes = [i, i, i, i]
    ll = ctypes.c_longlong
    lib.pyscfad_amd_ccsd_t_num_triples.restype = ll
    lib.pyscfad_amd_ccsd_t_num_triples.argtypes = [i]
    lib.pyscfad_amd_ccsd_t_energy_slice.restype = i
    lib.pyscfad_amd_ccsd_t_energy_slice.argtypes = [i, i] + [dp] * 7 + [i, i, ll, ll, dp, vp, sz, vp]
    lib.pyscfad_amd_ccsd_t_energy_vjp_slice.restype = i
    lib.pyscfad_amd_ccsd_t_energy_vjp_slice.argtypes = ([i, i] + [dp] * 7 + [i, d, i, ll, ll] + [dp] * 8
                                                        + [vp, sz, vp])
    ip = ctypes.POINTER(Instance)
    lib.pyscfad_amd_ccsd_t_batched_workspace_bytes.restype = sz
    lib.pyscfad_amd_ccsd_t_batched_workspace_bytes.argtypes = [i, ip, i, i]
    lib.pyscfad_amd_ccsd_t_batched.restype = i
    lib.pyscfad_amd_ccsd_t_batched.argtypes = [i, ip, i, i, vp, sz, vp]
    lib.pyscfad_amd_ccsd_t_autotune.restype = i
    lib.pyscfad_amd_ccsd_t_autotune.argtypes = [i, i, i, ctypes.POINTER(ctypes.c_int), vp]
    lib.pyscfad_amd_ccsd_t_df_workspace_bytes.restype = sz
    lib.pyscfad_amd_ccsd_t_df_workspace_bytes.argtypes = [i, i, i, i, i]
    lib.pyscfad_amd_ccsd_t_df_energy.restype = i
    lib.pyscfad_amd_ccsd_t_df_energy.argtypes = [i, i, i] + [dp] * 8 + [i, ll, ll, dp, vp, sz, vp]
    lib.pyscfad_amd_ccsd_t_df_energy_vjp.restype = i
    lib.pyscfad_amd_ccsd_t_df_energy_vjp.argtypes = ([i, i, i] + [dp] * 8 + [d, i, ll, ll] + [dp] * 9
                                                     + [vp, sz, vp])
    lib.pyscfad_amd_lno_ccsd_t_workspace_bytes.restype = sz
    lib.pyscfad_amd_lno_ccsd_t_workspace_bytes.argtypes = [i, i, i, i]
    lib.pyscfad_amd_lno_ccsd_t_energy.restype = i
    lib.pyscfad_amd_lno_ccsd_t_energy.argtypes = [i, i] + [dp] * 8 + [i, i, dp, vp, sz, vp]
    lib.pyscfad_amd_lno_ccsd_t_energy_vjp.restype = i
    lib.pyscfad_amd_lno_ccsd_t_energy_vjp.argtypes = [i, i] + [dp] * 8 + [i, d, i] + [dp] * 9 + [vp, sz, vp]
    lib.pyscfad_amd_last_error.restype = ctypes.c_char_p
    lib.pyscfad_amd_last_error.argtypes = []
    lib.pyscfad_amd_set_timing.restype = None
    lib.pyscfad_amd_set_timing.argtypes = [i]
    lib.pyscfad_amd_last_timing.restype = i
    lib.pyscfad_amd_last_timing.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong), i]
    for name in ("pyscfad_amd_xla_ccsd_t_energy", "pyscfad_amd_xla_ccsd_t_energy_vjp"):
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = [vp, ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, sz]
    lib.pyscfad_amd_graph_cache_stats.restype = None
    lib.pyscfad_amd_graph_cache_stats.argtypes = [ctypes.POINTER(ctypes.c_longlong)] * 2
    lib.pyscfad_amd_abi_version.restype = i
    lib.pyscfad_amd_abi_version.argtypes = []
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().pyscfad_amd_last_error().decode()
        raise T3Error(f"pyscfad_amd error {rc}: {msg}")


FAMILIES = ("gemm", "epilogue", "vpart", "reduce", "prep",
            "gemm_w", "gemm_b1", "gemm_b2", "gemm_b3", "gemm_b4", "df")


def set_timing(enable):
    load().pyscfad_amd_set_timing(int(bool(enable)))


def last_timing():
    """{family: (ms, launches)} of the calling thread's last call with timing on."""
    n = len(FAMILIES)
    ms = (ctypes.c_double * n)()
    nl = (ctypes.c_longlong * n)()
    m = load().pyscfad_amd_last_timing(ms, nl, n)
    return {FAMILIES[f]: (ms[f], nl[f]) for f in range(min(m, len(FAMILIES)))}


def graph_cache_stats():
    """(replays, captures) of the batched entry point's graph cache."""
    r, c = ctypes.c_longlong(0), ctypes.c_longlong(0)
    load().pyscfad_amd_graph_cache_stats(ctypes.byref(r), ctypes.byref(c))
    return r.value, c.value


class XlaOpaque(ctypes.Structure):
    """PyscfadAmdXlaOpaque of include/pyscfad_amd.h (the opaque bytes of the XLA custom calls)."""
    _fields_ = [("nocc", ctypes.c_int), ("nvir", ctypes.c_int), ("precision", ctypes.c_int),
                ("ovvv_packed", ctypes.c_int), ("workspace_bytes", ctypes.c_ulonglong)]
