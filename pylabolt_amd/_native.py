"""ctypes binding of libd2q9.so (include/d2q9.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already loaded (same soname): device pointers from torch
tensors and torch streams are then valid in the library.  There is no CPU
fallback: if the library is missing, or no GPU is visible, ``lib()`` raises.
"""
import ctypes
import os

import numpy as np
import torch  # noqa: F401  (must precede loading the HIP library)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libd2q9.so")

STORE_F64, STORE_F32, STORE_F16DEV, STORE_BF16DEV = 0, 1, 2, 3
BGK, MRT = 0, 1
FORCING_NONE, GUO_LINEAR, GUO_SECOND_ORDER = 0, 1, 2
FSRC_GRAVITY, FSRC_FIELD = 0, 1
CELL_FLUID, CELL_SOLID, CELL_NEAR_SOLID, CELL_BC, CELL_STORE_MOM = (
    1, 2, 4, 8, 16)
RULE_STREAM, RULE_BB, RULE_VEL, RULE_PRES, RULE_ZG = 0, 1, 2, 3, 4

STORAGE_CODES = {"f64": STORE_F64, "f32": STORE_F32,
                 "f16dev": STORE_F16DEV, "bf16dev": STORE_BF16DEV}
FORCING_CODES = {None: FORCING_NONE, "guo_linear": GUO_LINEAR,
                 "guo_second_order": GUO_SECOND_ORDER}
MODEL_CODES = {"BGK": BGK, "MRT": MRT}


class BcElem(ctypes.Structure):
    _fields_ = [("vec", ctypes.c_double * 2), ("scalar", ctypes.c_double),
                ("normal", ctypes.c_int32 * 2), ("type", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class BcNode(ctypes.Structure):
    _fields_ = [("rule", ctypes.c_uint16 * 9), ("pad", ctypes.c_uint16)]


class FConst(ctypes.Structure):
    """Mirror of ``d2q9_fconst``."""
    _fields_ = [("w", ctypes.c_float * 9), ("rates", ctypes.c_float * 9),
                ("scale", ctypes.c_float * 9), ("inv_scale", ctypes.c_float * 9),
                ("omega", ctypes.c_float), ("inv_cs_2", ctypes.c_float),
                ("inv_cs_4", ctypes.c_float), ("float_min", ctypes.c_float),
                ("gravity", ctypes.c_float * 2)]


class Lattice(ctypes.Structure):
    """Mirror of ``d2q9_lattice``."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("pitch", ctypes.c_int32), ("origin", ctypes.c_int32),
        ("plane", ctypes.c_int64),
        ("cpl", ctypes.c_int32), ("nbj", ctypes.c_int32),
        ("pop_len", ctypes.c_int64),
        ("storage", ctypes.c_int32), ("collision", ctypes.c_int32),
        ("forcing", ctypes.c_int32), ("force_source", ctypes.c_int32),
        ("x_periodic", ctypes.c_int32), ("y_periodic", ctypes.c_int32),
        ("tile_j", ctypes.c_int32), ("tile_i", ctypes.c_int32),
        ("omega", ctypes.c_double), ("inv_cs_2", ctypes.c_double),
        ("inv_cs_4", ctypes.c_double), ("float_min", ctypes.c_double),
        ("w", ctypes.c_double * 9), ("gravity", ctypes.c_double * 2),
        ("rates", ctypes.c_double * 9), ("dev_scale", ctypes.c_double * 9),
        ("cf", FConst),
        ("flags", ctypes.c_void_p), ("tiles", ctypes.c_void_p),
        ("n_bulk", ctypes.c_int32), ("n_general", ctypes.c_int32),
        ("fix_cells", ctypes.c_void_p),
        ("n_fix", ctypes.c_int32), ("pad_fix", ctypes.c_int32),
        ("bulk_masks", ctypes.c_void_p),
        ("bc_slot", ctypes.c_void_p), ("bc_nodes", ctypes.c_void_p),
        ("bc_elems", ctypes.c_void_p), ("force_field", ctypes.c_void_p),
        ("wall_index", ctypes.c_void_p), ("wall_masks", ctypes.c_void_p),
    ]


# numpy mirrors of d2q9_force_node (16 B) and d2q9_force_group (32 B)
FORCE_NODE_DTYPE = np.dtype([("cell", "<i8"), ("group", "<i4"),
                              ("mask", "<u4")])
FORCE_GROUP_DTYPE = np.dtype([("ref", "<f8", (2,)), ("torque", "<i4"),
                               ("first_block", "<i4"), ("n_blocks", "<i4"),
                               ("pad", "<i4")])
FORCE_BLOCK = 256
assert FORCE_NODE_DTYPE.itemsize == 16 and FORCE_GROUP_DTYPE.itemsize == 32


class BatchEntry(ctypes.Structure):
    """Mirror of ``d2q9_batch_entry``."""
    _fields_ = [("L", Lattice), ("f", ctypes.c_void_p * 2),
                ("mom", ctypes.c_void_p * 2)]


class BatchItem(ctypes.Structure):
    """Mirror of ``d2q9_batch_item``."""
    _fields_ = [("lattice", ctypes.c_int32), ("code", ctypes.c_int32)]


assert ctypes.sizeof(BcElem) == 40 and ctypes.sizeof(BcNode) == 20

_P = ctypes.c_void_p
_LP = ctypes.POINTER(Lattice)
# name -> (restype, argtypes)
SIGNATURES = {
    "d2q9_init_eq": (ctypes.c_int, [_LP, _P, _P, _P, _P]),
    "d2q9_collide": (ctypes.c_int, [_LP, _P, _P, _P, _P, ctypes.c_int, _P]),
    "d2q9_stream_bc": (ctypes.c_int, [_LP, _P, _P, _P, _P]),
    "d2q9_fused_step": (ctypes.c_int, [_LP, _P, _P, _P, _P, ctypes.c_int, _P]),
    "d2q9_fused_step2": (ctypes.c_int, [_LP, _LP, _LP, _P, _P,
                                        ctypes.c_int32, _P, _P, _P, _P, _P,
                                        _P, ctypes.c_int, _P]),
    "d2q9_deep_rows": (ctypes.c_int, []),
    "d2q9_moments": (ctypes.c_int, [_LP, _P, _P, _P]),
    "d2q9_reduce_work_len": (ctypes.c_int64, [_LP]),
    "d2q9_mass_fp64": (ctypes.c_int, [_LP, _P, _P, _P, _P]),
    "d2q9_mass_fp64_static": (ctypes.c_int, [_LP, _P, _P, _P]),
    "d2q9_residues": (ctypes.c_int, [_LP, _P, _P, _P, _P, _P]),
    "d2q9_copy": (ctypes.c_int, [_P, _P, ctypes.c_int64, _P]),
    "d2q9_graph_begin": (ctypes.c_int, [_P]),
    "d2q9_graph_end": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_void_p)]),
    "d2q9_graph_launch": (ctypes.c_int, [_P, _P]),
    "d2q9_graph_destroy": (ctypes.c_int, [_P]),
    "d2q9_last_error": (ctypes.c_char_p, []),
    "d2q9_source_hash": (ctypes.c_char_p, []),
    "d2q9_abi_layout": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int64),
                                       ctypes.c_int]),
    "d2q9_adjoint_step": (ctypes.c_int, [_LP, ctypes.c_int, _P, _P, _P, _P,
                                         _P, _P]),
    "d2q9_adjoint_init_grad": (ctypes.c_int, [_LP, _P, _P, _P, _P]),
    "d2q9_batched_step": (ctypes.c_int, [_P, _P, _P, _P] + [ctypes.c_int] * 9
                          + [_P]),
    "d2q9_momentum_exchange": (ctypes.c_int, [_LP, _P, _P, _P, _P,
                                              ctypes.c_int64, _P, ctypes.c_int,
                                              _P, _P, _P]),
}

# fields whose offsets d2q9_abi_layout reports, in its order
ABI_FIELDS = ["plane", "cpl", "pop_len", "storage", "tile_j", "omega", "w",
              "gravity", "rates", "dev_scale", "cf", "flags", "tiles",
              "n_bulk", "fix_cells", "n_fix", "bulk_masks", "bc_slot",
              "bc_nodes",
              "bc_elems", "force_field", "wall_index", "wall_masks"]


def abi_mismatches(handle):
    """Compare the ctypes mirror with the library's own struct layout."""
    out = (ctypes.c_int64 * 32)()
    n = handle.d2q9_abi_layout(out, 32)
    got = list(out[:n])
    want = [ctypes.sizeof(Lattice), ctypes.sizeof(BcElem),
            ctypes.sizeof(BcNode)]
    want += [getattr(Lattice, f).offset for f in ABI_FIELDS]
    want += [ctypes.sizeof(BatchEntry), BatchEntry.f.offset,
             BatchEntry.mom.offset, ctypes.sizeof(BatchItem)]
    return [(i, g, w) for i, (g, w) in enumerate(zip(got, want)) if g != w] \
        + ([("count", n, len(want))] if n != len(want) else [])

_lib = None


class NativeError(RuntimeError):
    pass


def load_library(path=LIB_PATH, csrc=None, header=None):
    """Load the library and declare every prototype (no GPU needed).

    Refuses a library built from other sources than those of the tree it is
    loaded from (``csrc`` and ``header`` default to the package's own): the
    digest the build stamped into it must equal the tree's."""
    from pylabolt_amd import build as _build
    if not os.path.exists(path):
        raise NativeError(
            f"{path} is missing: build it with `python -m pylabolt_amd.build`"
        )
    csrc = _build.CSRC if csrc is None else csrc
    header = _build.HEADER if header is None else header
    handle = ctypes.CDLL(path)
    if not hasattr(handle, "d2q9_source_hash"):
        raise NativeError(f"{path} carries no source digest: rebuild it")
    handle.d2q9_source_hash.restype = ctypes.c_char_p
    built = handle.d2q9_source_hash().decode()
    want = "h" + _build.source_digest(csrc, header)
    if built != want:
        raise NativeError(
            f"{path} is stale: built from sources with digest {built}, the "
            f"tree has {want}; rebuild with `python -m pylabolt_amd.build`")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    bad = abi_mismatches(handle)
    if bad:
        raise NativeError(f"d2q9_lattice layout differs from the library: {bad}")
    return handle


def lib():
    """The loaded library; raises if it is missing or no GPU is visible."""
    global _lib
    if _lib is None:
        if not torch.cuda.is_available():
            raise NativeError("no HIP device visible: the D2Q9 engine runs on "
                              "the GPU only")
        _lib = load_library()
    return _lib


def check(rc, what):
    """Raise ``NativeError`` with the library's message when rc != 0."""
    if rc != 0:
        msg = lib().d2q9_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    """Device address of a tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())
