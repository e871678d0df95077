"""ctypes wrapper over liboracle.so -- the CPU restatement of par2-rs's path.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker.  The product
(``par2-rs_amd/libpar2hip.so`` and its Python binding) never imports this.

Each wrapper names the C function it calls; the C source cites the reference
(par2-rs 0.3.1) file:line it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


class _FileIn(C.Structure):
    _fields_ = [("name", C.c_char_p), ("data", C.c_void_p), ("len", C.c_uint64)]


class _VolumeOut(C.Structure):
    _fields_ = [("name", C.c_char * 1024), ("bytes", C.c_void_p), ("len", C.c_uint64)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_library(path: str) -> None:
    """Load the restatement from another build of the same sources (for
    instance one compiled for the running host's CPU with -march=native);
    every wrapper below then calls into it."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None
    lib()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        S = C.c_size_t
        L.po_init.restype = None
        L.po_tables.argtypes = [P, P]
        L.po_gf_mul.argtypes = [C.c_uint16, C.c_uint16]
        L.po_gf_mul.restype = C.c_uint16
        L.po_gf_div.argtypes = [C.c_uint16, C.c_uint16, C.POINTER(C.c_uint16)]
        L.po_gf_div.restype = C.c_int
        L.po_gf_pow.argtypes = [C.c_uint16, C.c_uint64]
        L.po_gf_pow.restype = C.c_uint16
        L.po_par2_constants.argtypes = [S, P]
        L.po_coeff_matrix.argtypes = [P, S, P, S, P]
        L.po_mul_slice.argtypes = [C.c_uint16, P, S]
        L.po_muladd.argtypes = [P, P, S, C.c_uint16]
        L.po_muladd_region.argtypes = [P, S, P, S, P, S, S, S]
        L.po_encode.argtypes = [P, S, S, P, S, P]
        L.po_transform.argtypes = [P, P, S, P, S, P, P, P, P, P]
        L.po_transform.restype = C.c_int
        L.po_reconstruct_chunk.argtypes = [S, S, P, P, P, P, P, P, S, P]
        L.po_gf_invert.argtypes = [P, S, P]
        L.po_gf_invert.restype = C.c_int
        L.po_md5.argtypes = [P, S, P]
        L.po_crc32.argtypes = [P, S]
        L.po_crc32.restype = C.c_uint32
        L.po_crc32_update.argtypes = [C.c_uint32, P, S]
        L.po_crc32_update.restype = C.c_uint32
        L.po_hash_slices.argtypes = [P, S, S, S, P, P, P]
        L.po_encode_avx2.argtypes = [P, S, S, P, S, P, C.c_int]
        L.po_encode_avx2.restype = C.c_int
        L.po_hash_slices_mt.argtypes = [P, S, S, P, P, C.c_int]
        L.po_have_avx2.restype = C.c_int
        L.po_init()
        L.po_par2_block_size.argtypes = [C.c_uint64]
        L.po_par2_block_size.restype = C.c_uint64
        L.po_par2_file_id.argtypes = [P, C.c_uint64, C.c_char_p, P]
        L.po_par2_padding_width.argtypes = [C.c_uint32]
        L.po_par2_padding_width.restype = C.c_int
        L.po_par2_volume_name.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.c_int, C.c_char_p, S]
        L.po_par2_create.argtypes = [P, S, C.c_uint64, C.c_float, C.c_int, C.c_char_p,
                                     C.POINTER(C.POINTER(_VolumeOut)), C.POINTER(S)]
        L.po_par2_create.restype = C.c_int
        L.po_par2_free.argtypes = [C.POINTER(_VolumeOut), S]
        L.po_par2_free.restype = None
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


# ---- field (src/galois/core.rs) -------------------------------------------
def tables():
    log = np.zeros(65536, np.uint16)
    exp = np.zeros(131070, np.uint16)
    lib().po_tables(_p(log), _p(exp))
    return log, exp


def gf_mul(a: int, b: int) -> int:
    return int(lib().po_gf_mul(a, b))


def gf_div(a: int, b: int):
    q = C.c_uint16(0)
    return None if lib().po_gf_div(a, b, C.byref(q)) else int(q.value)


def gf_pow(a: int, n: int) -> int:
    return int(lib().po_gf_pow(a, n))


def par2_constants(count: int) -> np.ndarray:
    out = np.zeros(max(count, 1), np.uint16)
    lib().po_par2_constants(count, _p(out))
    return out[:count]


def coeff_matrix(n: int, exps) -> np.ndarray:
    exps = np.ascontiguousarray(exps, np.uint32)
    consts = par2_constants(n)
    out = np.zeros((len(exps), n), np.uint16)
    lib().po_coeff_matrix(_p(consts), n, _p(exps), len(exps), _p(out))
    return out


# ---- scalar strategy (src/galois/scalar.rs) ------------------------------
def mul_slice(k: int, data: np.ndarray) -> None:
    lib().po_mul_slice(k, _p(data), data.size)


def muladd(dst: np.ndarray, src: np.ndarray, k: int) -> None:
    lib().po_muladd(_p(dst), _p(src), min(dst.size, src.size), k)


def muladd_region(dsts, srcs, coeff: np.ndarray, off: int, length: int) -> None:
    coeff = np.ascontiguousarray(coeff, np.uint16)
    lib().po_muladd_region(_ptrs(dsts), len(dsts), _ptrs(srcs), len(srcs), _p(coeff),
                           coeff.shape[1], off, length)


def muladd_region_rows(dsts, srcs, rows, off: int, length: int) -> None:
    """gf_muladd_region_scalar (src/galois/scalar.rs:82-113) with coefficient
    rows of any length: row i applies entries j < min(len(rows[i]), nsrc),
    the `.enumerate().take(num_srcs)` of scalar.rs:100."""
    for d, row in zip(dsts, rows):
        row = np.ascontiguousarray(row, np.uint16).reshape(1, -1)
        k = min(row.shape[1], len(srcs))
        if k:
            muladd_region([d], srcs[:k], row, off, length)


# ---- codec (src/par2_rs.rs) ------------------------------------------------
def encode(data: np.ndarray, exps) -> np.ndarray:
    """data: (n, words) uint16 -> (r, words) recovery slices."""
    data = np.ascontiguousarray(data, np.uint16)
    exps = np.ascontiguousarray(exps, np.uint32)
    n, words = data.shape
    rec = np.zeros((len(exps), words), np.uint16)
    lib().po_encode(_p(data), n, words, _p(exps), len(exps), _p(rec))
    return rec


def encode_avx2(data: np.ndarray, exps, threads: int = 0) -> np.ndarray:
    data = np.ascontiguousarray(data, np.uint16)
    exps = np.ascontiguousarray(exps, np.uint32)
    n, words = data.shape
    rec = np.zeros((len(exps), words), np.uint16)
    if lib().po_encode_avx2(_p(data), n, words, _p(exps), len(exps), _p(rec), threads) != 0:
        raise RuntimeError("AVX2 unavailable")
    return rec


def have_avx2() -> bool:
    return bool(lib().po_have_avx2())


def transform(n_total: int, missing, present, rec_exps):
    """Faithful ReconstructionTransform (par2_rs.rs:319-422)."""
    consts = par2_constants(n_total)
    missing = np.ascontiguousarray(missing, np.uintp)
    present = np.ascontiguousarray(present, np.uintp)
    rec_exps = np.ascontiguousarray(rec_exps, np.uint32)
    m = len(missing)
    row_order = np.zeros(m, np.uintp)
    scale = np.zeros(m, np.uint16)
    elim = np.zeros((m, m), np.uint16)
    vand = np.zeros((m, max(len(present), 1)), np.uint16)
    rc = lib().po_transform(_p(consts), _p(missing), m, _p(present), len(present), _p(rec_exps),
                            _p(row_order), _p(scale), _p(elim), _p(vand))
    if rc == -1:
        raise ValueError("Singular matrix")
    if rc == -2:
        raise ValueError("No missing blocks to reconstruct")
    return row_order, scale, elim, vand[:, : len(present)]


def reconstruct(n_total: int, missing, present, rec_exps, rec_slices, data_slices) -> np.ndarray:
    """Rebuild the missing slices with the reference's transform + replay
    (par2_rs.rs:442-675).  rec_slices: (m, words); data_slices: (len(present), words)."""
    row_order, scale, elim, vand = transform(n_total, missing, present, rec_exps)
    m = len(missing)
    rec_slices = [np.ascontiguousarray(r, np.uint16) for r in rec_slices]
    data_slices = [np.ascontiguousarray(d, np.uint16) for d in data_slices]
    words = rec_slices[0].size
    out = [np.zeros(words, np.uint16) for _ in range(m)]
    vand = np.ascontiguousarray(vand)
    lib().po_reconstruct_chunk(m, len(data_slices), _p(row_order), _p(scale), _p(elim), _p(vand),
                               _ptrs(rec_slices), _ptrs(data_slices) if data_slices else None,
                               words, _ptrs(out))
    return np.stack(out), row_order


def gf_invert(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint16)
    m = a.shape[0]
    inv = np.zeros_like(a)
    if lib().po_gf_invert(_p(a), m, _p(inv)) != 0:
        raise ValueError("Singular matrix")
    return inv


# ---- hashes ------------------------------------------------------------------
def md5(data: bytes) -> bytes:
    buf = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(16, np.uint8)
    lib().po_md5(_p(buf), len(data), _p(out))
    return out.tobytes()


def crc32(data: bytes) -> int:
    buf = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
    return int(lib().po_crc32(_p(buf), len(data)))


def hash_slices(data: np.ndarray, n: int, slice_size: int, slice_stride: int | None = None, lens=None):
    """IFSC MD5/CRC32 of n zero-padded slices (creator.rs:389-409)."""
    data = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    md5s = np.zeros((max(n, 1), 16), np.uint8)
    crcs = np.zeros(max(n, 1), np.uint32)
    lens_a = None if lens is None else np.ascontiguousarray(lens, np.uintp)
    lib().po_hash_slices(_p(data), n, slice_size, slice_stride or slice_size,
                         None if lens_a is None else _p(lens_a), _p(md5s), _p(crcs))
    return md5s[:n], crcs[:n]


def hash_slices_mt(data: np.ndarray, n: int, slice_size: int, threads: int = 0):
    data = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    md5s = np.zeros((n, 16), np.uint8)
    crcs = np.zeros(n, np.uint32)
    lib().po_hash_slices_mt(_p(data), n, slice_size, _p(md5s), _p(crcs), threads)
    return md5s, crcs


# ---- PAR2 creation (src/creator.rs:261-531, src/writer.rs, src/volumes.rs) ----
SCHEME_SINGLE, SCHEME_EXPONENTIAL = 0, 1


def par2_block_size(total: int) -> int:
    """calculate_block_size (creator.rs:29-67), po_par2_block_size."""
    return int(lib().po_par2_block_size(total))


def par2_file_id(hash16k: bytes, length: int, name: str) -> bytes:
    h = np.frombuffer(hash16k, np.uint8).copy()
    out = np.zeros(16, np.uint8)
    lib().po_par2_file_id(_p(h), length, name.encode(), _p(out))
    return out.tobytes()


def par2_volume_name(stem: str, start: int, end: int, width: int) -> str:
    buf = C.create_string_buffer(1024)
    lib().po_par2_volume_name(stem.encode(), start, end, width, buf, 1024)
    return buf.value.decode()


def par2_padding_width(max_exponent: int) -> int:
    return int(lib().po_par2_padding_width(max_exponent))


def par2_create(files, block_size: int = 0, redundancy: float = 5.0, scheme: int = SCHEME_EXPONENTIAL,
                stem: str = "out"):
    """Par2Creator::create on in-memory files [(name, bytes)]: list of
    (output path, bytes), the main .par2 first (po_par2_create)."""
    bufs = [np.frombuffer(d, np.uint8) if len(d) else np.zeros(1, np.uint8) for _, d in files]
    arr = (_FileIn * len(files))()
    names = [n.encode() for n, _ in files]
    for i, ((_, d), b) in enumerate(zip(files, bufs)):
        arr[i].name = names[i]
        arr[i].data = b.ctypes.data
        arr[i].len = len(d)
    vols = C.POINTER(_VolumeOut)()
    nv = C.c_size_t(0)
    rc = lib().po_par2_create(arr, len(files), block_size, redundancy, scheme, stem.encode(), C.byref(vols),
                              C.byref(nv))
    if rc != 0:
        raise ValueError(f"po_par2_create: {rc}")
    out = [(vols[i].name.decode(), C.string_at(vols[i].bytes, vols[i].len)) for i in range(nv.value)]
    lib().po_par2_free(vols, nv.value)
    return out
