"""Python binding of libpar2hip.so (the C ABI in include/par2hip.h).

Mirrors the reference's plugin interface for the Reed-Solomon hot path:

* :class:`HipStrategy` -- the methods of ``GaloisSimdStrategy``
  (src/galois/simd/mod.rs:24-128): ``name``, ``is_available``, ``priority``,
  ``mul_slice``, ``muladd``, ``muladd_region``, plus the shuffle2x hooks, which
  are identity/fallbacks here (``supports_shuffle2x`` is False, as the GPU
  consumes the interleaved layout).
* :class:`Context` -- the coarse whole-set hooks: :meth:`Context.encode`
  (StreamingEncoder, src/par2_rs.rs:687-902), :meth:`Context.reconstruct`
  (compute_reconstruction_transform + reconstruct_streaming_chunk,
  src/par2_rs.rs:319-675), :meth:`Context.hash_slices` (IFSC,
  src/creator.rs:397-409), :meth:`Context.gf_invert`, and device-resident
  variants for callers that keep sets in HBM.

Errors surface as the reference's error kinds (src/error.rs:5-39).  The
library is required: importing succeeds without it, but every compute call
raises if ``libpar2hip.so`` is missing -- there is no CPU fallback here.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpar2hip.so")

OK, EINVAL, ENODEV, ENOMEM, ESINGULAR, EHIP, ERECOVERY, EREPAIR, ECANCELLED, ETIMEDOUT = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9)
PRIORITY_OPTIMAL = 40  # Priority::Optimal (src/galois/simd/mod.rs:20)


class Par2Error(Exception):
    """Base error (src/error.rs)."""


class InvalidFormat(Par2Error):
    pass


class RepairFailed(Par2Error):
    pass


class InsufficientRecovery(Par2Error):
    pass


class DeviceError(Par2Error):
    pass


class Cancelled(Par2Error):
    """Par2Error::Cancelled (src/error.rs:35)."""


JOB_ENCODE, JOB_REPAIR = 0, 1


class Job(C.Structure):
    """par2hip_job: one PAR2 set to encode or repair in a batch."""
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("d_slices", C.c_void_p), ("n", C.c_uint64),
                ("slice_size", C.c_uint64), ("exps", C.c_void_p), ("count", C.c_uint64), ("d_recovery", C.c_void_p),
                ("missing", C.c_void_p), ("d_out", C.c_void_p), ("d_md5", C.c_void_p), ("d_crc", C.c_void_p)]


class Request(C.Structure):
    """par2hip_request: one request to the serving loop."""
    _fields_ = [("job", Job), ("hashes", C.c_int32), ("reserved", C.c_int32), ("arrival_ms", C.c_double)]


class Result(C.Structure):
    """par2hip_result: a finished request (outputs stay until release)."""
    _fields_ = [("status", C.c_int32), ("batch_size", C.c_int32), ("d_out", C.c_void_p), ("d_md5", C.c_void_p),
                ("d_crc", C.c_void_p), ("t_arrival_ms", C.c_double), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_start_ms", C.c_double), ("t_done_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "failed", "batches", "max_batch_seen",
                                           "charged", "peak_charged", "pending", "admitted", "arena_fallbacks")]


class InputFile(C.Structure):
    _fields_ = [("name", C.c_char_p), ("data", C.c_void_p), ("len", C.c_uint64)]


class OutputFile(C.Structure):
    _fields_ = [("path", C.c_char * 1024), ("bytes", C.c_uint64)]


VOLUMES_SINGLE, VOLUMES_EXPONENTIAL = 0, 1


class SetInfo(C.Structure):
    _fields_ = [("block_size", C.c_uint64), ("blocks", C.c_uint64), ("files", C.c_uint32),
                ("recovery_slices", C.c_uint32), ("min_exponent", C.c_uint32), ("max_exponent", C.c_uint32),
                ("set_id", C.c_uint8 * 16)]


class RepairReport(C.Structure):
    _fields_ = [("block_size", C.c_uint64), ("blocks", C.c_uint64), ("blocks_damaged", C.c_uint64),
                ("files", C.c_uint32), ("files_damaged", C.c_uint32), ("files_missing", C.c_uint32),
                ("recovery_slices", C.c_uint32), ("repaired", C.c_int32), ("reserved", C.c_int32)]


class DeviceInfo(C.Structure):
    _fields_ = [("compute_units", C.c_int32), ("hash_cus", C.c_int32), ("arch", C.c_char * 32)]


class PlanTuning(C.Structure):
    _fields_ = [("policy", C.c_int32), ("policies", C.c_int32), ("ms", C.c_double * 4)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


_lib = None
_lib_lock = threading.Lock()


def build() -> None:
    """Compile libpar2hip.so for gfx950 in-tree."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc")], check=True)


def lib():
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise DeviceError(f"{LIB_PATH} is not built (run par2hip.build())")
            L = C.CDLL(LIB_PATH)
            P, S, I32 = C.c_void_p, C.c_size_t, C.c_int32
            sig = {
                "par2hip_available": ([], I32),
                "par2hip_strategy_name": ([], C.c_char_p),
                "par2hip_strategy_priority": ([], I32),
                "par2hip_ctx_create": ([I32, C.POINTER(P)], I32),
                "par2hip_ctx_destroy": ([P], I32),
                "par2hip_last_error": ([P], C.c_char_p),
                "par2hip_dev_alloc": ([P, S, C.POINTER(P)], I32),
                "par2hip_dev_free": ([P, P], I32),
                "par2hip_h2d": ([P, P, P, S], I32),
                "par2hip_d2h": ([P, P, P, S], I32),
                "par2hip_d2d": ([P, P, P, S], I32),
                "par2hip_sync": ([P], I32),
                "par2hip_gf_mul_slice": ([P, C.c_uint16, P, S], I32),
                "par2hip_gf_muladd": ([P, P, P, S, C.c_uint16], I32),
                "par2hip_gf_muladd_region": ([P, P, S, P, S, P, P, S, S], I32),
                "par2hip_encode": ([P, P, S, S, P, S, P, P, P], I32),
                "par2hip_encode_dev": ([P, P, S, S, P, S, P, P, P], I32),
                "par2hip_encode_dev_noalloc": ([P, P, S, S, P, S, P, P, P], I32),
                "par2hip_gf_invert": ([P, P, S, P], I32),
                "par2hip_reconstruct": ([P, P, S, S, P, S, P, P, P], I32),
                "par2hip_reconstruct_dev": ([P, P, S, S, P, S, P, P, P, I32], I32),
                "par2hip_repair_set_dev": ([P, P, S, S, P, P, P, P, S, P, P, C.POINTER(S)], I32),
                "par2hip_batch_run_dev": ([P, P, S], I32),
                "par2hip_get_device_info": ([P, P], I32),
                "par2hip_hash_slices": ([P, P, S, S, S, P, P, P], I32),
                "par2hip_hash_slices_dev": ([P, P, S, S, S, P, P, P], I32),
                "par2hip_set_gf_kernel": ([P, I32], I32),
                "par2hip_profile_enable": ([P, I32], I32),
                "par2hip_profile_read": ([P, P, S, C.POINTER(S)], I32),
                "par2hip_create_par2": ([P, P, S, C.c_uint64, C.c_float, I32, C.c_char_p, P, S, C.POINTER(S)], I32),
                "par2hip_par2_info": ([C.c_char_p, P], I32),
                "par2hip_repair_files": ([P, C.c_char_p, C.c_char_p, I32, P], I32),
                "par2hip_repair_files_cancellable": ([P, C.c_char_p, C.c_char_p, I32, P, P], I32),
                "par2hip_set_autotune": ([P, I32], I32),
                "par2hip_force_tile_shape": ([P, I32, I32, I32], I32),
                "par2hip_autotune_save": ([C.c_char_p], I32),
                "par2hip_autotune_load": ([C.c_char_p, C.POINTER(S)], I32),
                "par2hip_batch_plan_create": ([P, P, S, I32, C.POINTER(P)], I32),
                "par2hip_batch_plan_launch": ([P], I32),
                "par2hip_batch_plan_destroy": ([P], I32),
                "par2hip_batch_plan_tuning_info": ([P, P], I32),
                "par2hip_server_create": ([I32, C.c_uint64, I32, I32, C.POINTER(P)], I32),
                "par2hip_server_destroy": ([P], I32),
                "par2hip_server_now_ms": ([P], C.c_double),
                "par2hip_server_submit": ([P, P, C.POINTER(C.c_uint64)], I32),
                "par2hip_server_wait": ([P, C.c_uint64, P], I32),
                "par2hip_server_wait_timeout": ([P, C.c_uint64, C.c_double, P], I32),
                "par2hip_server_release": ([P, C.c_uint64], I32),
                "par2hip_server_get_stats": ([P, P], I32),
            }
            for name, (args, res) in sig.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
            _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def available() -> int:
    """Number of gfx950 devices (GaloisSimdStrategy::is_available)."""
    try:
        return int(lib().par2hip_available())
    except (DeviceError, OSError):
        return 0


class DeviceBuffer:
    """Device memory owned by a Context's device (freed with the buffer)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        ptr = C.c_void_p()
        ctx._check(lib().par2hip_dev_alloc(ctx.handle, self.nbytes, C.byref(ptr)))
        self.ptr = ptr

    @property
    def addr(self) -> int:
        return int(self.ptr.value)

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise InvalidFormat("upload beyond buffer")
        self.ctx._check(lib().par2hip_h2d(self.ctx.handle, C.c_void_p(self.addr + offset), _p(arr), arr.nbytes))

    def download(self, nbytes: int | None = None, offset: int = 0) -> np.ndarray:
        nbytes = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(nbytes, np.uint8)
        self.ctx._check(lib().par2hip_d2h(self.ctx.handle, _p(out), C.c_void_p(self.addr + offset), nbytes))
        return out

    def view(self, offset: int) -> "DeviceView":
        """A non-owning pointer `offset` bytes into this buffer."""
        if not 0 <= offset <= self.nbytes:
            raise InvalidFormat("view beyond buffer")
        return DeviceView(self, offset)

    def free(self) -> None:
        if self.ptr is not None and self.ptr.value:
            lib().par2hip_dev_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if self.ctx.handle is not None:
                self.free()
        except Exception:
            pass


class DeviceView:
    """Borrowed device pointer into a DeviceBuffer (keeps the owner alive)."""

    def __init__(self, owner: DeviceBuffer, offset: int):
        self.owner = owner
        self.nbytes = owner.nbytes - offset
        self.ptr = C.c_void_p(owner.addr + offset)


class CancelFlag:
    """A one-byte flag shared with the library (the layout of a Rust
    AtomicBool); set() from any thread requests cancellation."""

    def __init__(self, value: bool = False):
        self.flag = C.c_uint8(1 if value else 0)

    def set(self) -> None:
        self.flag.value = 1

    def is_set(self) -> bool:
        return bool(self.flag.value)


class Context:
    """One device; wraps par2hip_ctx (calls on one Context are serialised)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().par2hip_ctx_create(device, C.byref(h))
        if rc != OK:
            raise DeviceError(f"par2hip_ctx_create failed ({rc}): no gfx950 device {device}")
        self.handle = h

    def close(self) -> None:
        if self.handle is not None:
            lib().par2hip_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc == OK:
            return
        msg = (lib().par2hip_last_error(self.handle) or b"").decode(errors="replace")
        if rc == ESINGULAR:
            raise RepairFailed(msg or "Singular matrix")
        if rc == ERECOVERY:
            raise InsufficientRecovery(msg)
        if rc == EINVAL:
            raise InvalidFormat(msg or "invalid argument")
        raise DeviceError(f"{rc}: {msg}")

    def device_info(self) -> dict:
        d = DeviceInfo()
        self._check(lib().par2hip_get_device_info(self.handle, C.byref(d)))
        return {"compute_units": d.compute_units, "hash_cus": d.hash_cus, "arch": d.arch.decode()}

    # ---- memory
    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def sync(self) -> None:
        self._check(lib().par2hip_sync(self.handle))

    # ---- coarse hooks
    def encode(self, slices: np.ndarray, exponents, hashes: bool = False):
        """slices: (n, slice_size) uint8, zero-padded.  Returns recovery (r, slice_size)
        [, md5 (n, 16), crc32 (n,)]."""
        slices = np.ascontiguousarray(slices, np.uint8)
        n, s = slices.shape
        exps = np.ascontiguousarray(exponents, np.uint32)
        r = len(exps)
        rec = np.zeros((max(r, 1), s), np.uint8)
        md5 = np.zeros((n, 16), np.uint8) if hashes else None
        crc = np.zeros(n, np.uint32) if hashes else None
        self._check(lib().par2hip_encode(self.handle, _p(slices), n, s, _p(exps), r, _p(rec), _p(md5), _p(crc)))
        rec = rec[:r]
        return (rec, md5, crc) if hashes else rec

    def encode_dev(self, d_slices: DeviceBuffer, n: int, slice_size: int, exponents, d_rec: DeviceBuffer,
                   d_md5: DeviceBuffer | None = None, d_crc: DeviceBuffer | None = None) -> None:
        exps = np.ascontiguousarray(exponents, np.uint32)
        self._check(lib().par2hip_encode_dev(self.handle, d_slices.ptr, n, slice_size, _p(exps), len(exps),
                                             d_rec.ptr, d_md5.ptr if d_md5 else None,
                                             d_crc.ptr if d_crc else None))

    def encode_dev_noalloc(self, d_slices: DeviceBuffer, n: int, slice_size: int, exponents, d_rec: DeviceBuffer,
                           d_md5: DeviceBuffer | None = None, d_crc: DeviceBuffer | None = None) -> None:
        """encode_dev through par2hip_encode_dev_noalloc (no allocation during
        the call; working memory in the library's per-device array)."""
        exps = np.ascontiguousarray(exponents, np.uint32)
        self._check(lib().par2hip_encode_dev_noalloc(self.handle, d_slices.ptr, n, slice_size, _p(exps), len(exps),
                                                     d_rec.ptr if d_rec else None, d_md5.ptr if d_md5 else None,
                                                     d_crc.ptr if d_crc else None))

    def reconstruct(self, slices: np.ndarray, missing, recovery: np.ndarray, rec_exps) -> np.ndarray:
        slices = np.ascontiguousarray(slices, np.uint8)
        n, s = slices.shape
        missing = np.ascontiguousarray(missing, np.uint32)
        m = len(missing)
        recovery = np.ascontiguousarray(recovery, np.uint8).reshape(-1, s)
        rec_exps = np.ascontiguousarray(rec_exps, np.uint32)
        if m == 0:
            return np.zeros((0, s), np.uint8)
        if recovery.shape[0] < m or len(rec_exps) < m:
            raise InsufficientRecovery(f"Need {m} recovery blocks, have {min(recovery.shape[0], len(rec_exps))}")
        out = np.zeros((m, s), np.uint8)
        self._check(lib().par2hip_reconstruct(self.handle, _p(slices), n, s, _p(missing), m,
                                              _p(np.ascontiguousarray(recovery[:m])), _p(rec_exps[:m].copy()),
                                              _p(out)))
        return out

    def reconstruct_dev(self, d_slices: DeviceBuffer, n: int, slice_size: int, missing,
                        d_recovery: DeviceBuffer, rec_exps, d_out: DeviceBuffer | None = None,
                        in_place: bool = False) -> None:
        missing = np.ascontiguousarray(missing, np.uint32)
        rec_exps = np.ascontiguousarray(rec_exps, np.uint32)
        self._check(lib().par2hip_reconstruct_dev(self.handle, d_slices.ptr, n, slice_size, _p(missing),
                                                  len(missing), d_recovery.ptr, _p(rec_exps),
                                                  d_out.ptr if d_out else None, 1 if in_place else 0))

    def repair_set_dev(self, d_slices: DeviceBuffer, n: int, slice_size: int, expected_md5: np.ndarray,
                       expected_crc: np.ndarray, d_recovery: DeviceBuffer | None, rec_exps,
                       d_out: DeviceBuffer | None = None):
        """Verify n device slices against their IFSC entries and rebuild the damaged
        ones (in place unless d_out is given).  Returns the damaged indices."""
        md5 = np.ascontiguousarray(expected_md5, np.uint8).reshape(n, 16)
        crc = np.ascontiguousarray(expected_crc, np.uint32)
        exps = np.ascontiguousarray(rec_exps if rec_exps is not None else [], np.uint32)
        dmg = np.zeros(max(n, 1), np.uint32)
        cnt = C.c_size_t(0)
        rc = lib().par2hip_repair_set_dev(self.handle, d_slices.ptr, n, slice_size, _p(md5), _p(crc),
                                          d_recovery.ptr if d_recovery else None, _p(exps) if len(exps) else None,
                                          len(exps), d_out.ptr if d_out else None, _p(dmg), C.byref(cnt))
        if rc == ERECOVERY:
            raise InsufficientRecovery(f"Need {cnt.value} recovery blocks, have {len(exps)}")
        self._check(rc)
        return dmg[:cnt.value].copy()

    def batch_run_dev(self, jobs) -> None:
        """Run many sets at once.  Each job is a dict: kind ('encode'|'repair'),
        slices (DeviceBuffer), n, slice_size, exps, recovery (DeviceBuffer),
        [missing, out (DeviceBuffer) for repair], [md5, crc (DeviceBuffer)]."""
        arr, keep = self._jobs(jobs)
        self._check(lib().par2hip_batch_run_dev(self.handle, arr, len(jobs)))

    def batch_plan(self, jobs, graph: bool = True) -> "BatchPlan":
        """Plan the jobs once (see batch_run_dev) for repeated launches,
        replayed as one HIP graph when graph is True."""
        arr, keep = self._jobs(jobs)
        h = C.c_void_p()
        self._check(lib().par2hip_batch_plan_create(self.handle, arr, len(jobs), 1 if graph else 0, C.byref(h)))
        return BatchPlan(self, h, jobs)

    def _jobs(self, jobs):
        arr = (Job * len(jobs))()
        keep = []
        for i, j in enumerate(jobs):
            e = np.ascontiguousarray(j.get("exps", []), np.uint32)
            miss = np.ascontiguousarray(j.get("missing", []), np.uint32)
            keep += [e, miss]
            a = arr[i]
            a.kind = JOB_REPAIR if j["kind"] == "repair" else JOB_ENCODE
            a.d_slices = j["slices"].ptr
            a.n = j["n"]
            a.slice_size = j["slice_size"]
            a.exps = e.ctypes.data if len(e) else None
            a.count = len(miss) if j["kind"] == "repair" else len(e)
            a.d_recovery = j["recovery"].ptr if j.get("recovery") is not None else None
            a.missing = miss.ctypes.data if len(miss) else None
            a.d_out = j["out"].ptr if j.get("out") is not None else None
            a.d_md5 = j["md5"].ptr if j.get("md5") is not None else None
            a.d_crc = j["crc"].ptr if j.get("crc") is not None else None
        return arr, keep

    def gf_invert(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, np.uint16)
        m = a.shape[0]
        inv = np.zeros_like(a)
        self._check(lib().par2hip_gf_invert(self.handle, _p(a), m, _p(inv)))
        return inv

    def hash_slices(self, data: np.ndarray, n: int, slice_size: int, slice_stride: int = 0, lens=None):
        data = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        md5 = np.zeros((max(n, 1), 16), np.uint8)
        crc = np.zeros(max(n, 1), np.uint32)
        lv = None if lens is None else np.ascontiguousarray(lens, np.uint64)
        self._check(lib().par2hip_hash_slices(self.handle, _p(data), n, slice_size, slice_stride, _p(lv),
                                              _p(md5), _p(crc)))
        return md5[:n], crc[:n]

    def hash_slices_dev(self, d_data: DeviceBuffer, n: int, slice_size: int, d_md5: DeviceBuffer | None,
                        d_crc: DeviceBuffer | None, slice_stride: int = 0, lens=None) -> None:
        lv = None if lens is None else np.ascontiguousarray(lens, np.uint64)
        self._check(lib().par2hip_hash_slices_dev(self.handle, d_data.ptr, n, slice_size, slice_stride, _p(lv),
                                                  d_md5.ptr if d_md5 else None, d_crc.ptr if d_crc else None))

    # ---- tuning / measurement
    def set_gf_kernel(self, which: int) -> int:
        rc = lib().par2hip_set_gf_kernel(self.handle, which)
        if rc < 0:
            self._check(rc)
        return rc

    def create_par2(self, files, out_stem: str, block_size: int = 0, redundancy: float = 5.0,
                    scheme: int = VOLUMES_EXPONENTIAL) -> list:
        """Par2Creator::create: files [(name, bytes)] -> written [(path, size)],
        the main .par2 first (par2hip_create_par2)."""
        arr = (InputFile * len(files))()
        keep = []
        for i, (name, data) in enumerate(files):
            b = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
            keep.append((name.encode(), b))
            arr[i].name = keep[-1][0]
            arr[i].data = b.ctypes.data
            arr[i].len = len(data)
        cap = 64
        while True:
            outs = (OutputFile * cap)()
            n = C.c_size_t(0)
            rc = lib().par2hip_create_par2(self.handle, arr, len(files), block_size, redundancy, scheme,
                                           out_stem.encode(), outs, cap, C.byref(n))
            if rc == EINVAL and n.value > cap:
                cap = n.value
                continue
            self._check(rc)
            return [(outs[i].path.decode(), int(outs[i].bytes)) for i in range(n.value)]

    def repair_files(self, par2_path: str, base_dir: str | None = None, verify_only: bool = False,
                     cancel: "CancelFlag | None" = None) -> dict:
        """Par2Repairer: verify the set's files and repair them in place through
        temp files (par2hip_repair_files); returns the report.  With a
        CancelFlag this is repair_with_callbacks_cancellable (src/lib.rs:196):
        a set flag makes the call raise Cancelled with every original intact."""
        rep = RepairReport()
        path_b, base_b = par2_path.encode(), base_dir.encode() if base_dir else None
        if cancel is None:
            rc = lib().par2hip_repair_files(self.handle, path_b, base_b, 1 if verify_only else 0, C.byref(rep))
        else:
            rc = lib().par2hip_repair_files_cancellable(self.handle, path_b, base_b, 1 if verify_only else 0,
                                                        C.addressof(cancel.flag), C.byref(rep))
        if rc == EREPAIR:
            raise RepairFailed((lib().par2hip_last_error(self.handle) or b"").decode())
        if rc == ECANCELLED:
            raise Cancelled("cancelled")
        self._check(rc)
        return {k: int(getattr(rep, k)) for k, _ in RepairReport._fields_ if k != "reserved"}

    def copy_dev(self, dst: DeviceBuffer, src: DeviceBuffer, nbytes: int) -> None:
        """Device-to-device copy on the ctx stream."""
        if nbytes > min(dst.nbytes, src.nbytes):
            raise InvalidFormat("copy beyond buffer")
        self._check(lib().par2hip_d2d(self.handle, dst.ptr, src.ptr, nbytes))

    def set_autotune(self, on: bool) -> bool:
        """Tile-shape autotuning on first use of each shape class; returns the previous setting."""
        rc = lib().par2hip_set_autotune(self.handle, 1 if on else 0)
        if rc < 0:
            self._check(rc)
        return bool(rc)

    TILE_SHAPES = [(8, 2, 16), (8, 1, 16), (8, 2, 32), (4, 4, 16), (4, 2, 16), (4, 1, 16), (4, 2, 32),
                   (2, 4, 16), (2, 2, 16), (2, 1, 16), (1, 4, 16), (1, 2, 16), (1, 1, 16)]

    def force_tile_shape(self, shape=None) -> None:
        """Use tile shape (v, g, d) for every LDS contraction; None clears it."""
        v, g, d = shape if shape else (0, 0, 0)
        self._check(lib().par2hip_force_tile_shape(self.handle, v, g, d))

    def profile(self, on: bool = True) -> None:
        self._check(lib().par2hip_profile_enable(self.handle, 1 if on else 0))

    def profile_read(self) -> dict:
        arr = (KernelStat * 64)()
        cnt = C.c_size_t(0)
        self._check(lib().par2hip_profile_read(self.handle, arr, 64, C.byref(cnt)))
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms))
                for i in range(min(cnt.value, 64))}


class BatchPlan:
    """A planned batch (Context.batch_plan); launch() queues one run."""

    def __init__(self, ctx: "Context", handle, jobs):
        self.ctx, self.handle, self.jobs = ctx, handle, jobs

    def launch(self) -> None:
        self.ctx._check(lib().par2hip_batch_plan_launch(self.handle))

    def tuning(self) -> dict:
        """The tile-shape policy the plan uses and, when it was tuned, the
        timed launch of every policy (ms)."""
        t = PlanTuning()
        self.ctx._check(lib().par2hip_batch_plan_tuning_info(self.handle, C.byref(t)))
        return {"policy": int(t.policy), "policies": int(t.policies), "ms": [round(float(x), 3) for x in t.ms]}

    def close(self) -> None:
        if self.handle is not None:
            lib().par2hip_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def par2_info(path: str) -> dict:
    """What a .par2 index and the volumes beside it hold (no device needed)."""
    info = SetInfo()
    _status(lib().par2hip_par2_info(path.encode(), C.byref(info)), f"par2_info {path}")
    d = {k: int(getattr(info, k)) for k, _ in SetInfo._fields_ if k != "set_id"}
    d["set_id"] = bytes(info.set_id)
    return d


def autotune_save(path: str) -> None:
    _status(lib().par2hip_autotune_save(path.encode()), f"autotune_save {path}")


def autotune_load(path: str) -> int:
    n = C.c_size_t(0)
    _status(lib().par2hip_autotune_load(path.encode(), C.byref(n)), f"autotune_load {path}")
    return n.value


def _status(rc: int, what: str) -> None:
    if rc == OK:
        return
    if rc == ESINGULAR:
        raise RepairFailed(what)
    if rc == ERECOVERY:
        raise InsufficientRecovery(what)
    if rc == EREPAIR:
        raise RepairFailed(what)
    if rc == EINVAL:
        raise InvalidFormat(what)
    raise DeviceError(f"{rc}: {what}")


class Server:
    """The serving loop (continuous batching, HBM-budget admission).

    submit() takes a request dict like Context.batch_run_dev's jobs (slices,
    n, slice_size, exps, and for repairs recovery + missing) without output
    buffers; the server allocates them.  wait() returns a dict with the
    device outputs (valid until release()) and the server-clock timestamps."""

    def __init__(self, device: int = 0, hbm_budget: int = 1 << 30, lanes: int = 2, max_batch: int = 256):
        h = C.c_void_p()
        _status(lib().par2hip_server_create(device, int(hbm_budget), lanes, max_batch, C.byref(h)), "server_create")
        self.handle = h
        self._keep: dict[int, tuple] = {}

    def close(self) -> None:
        if self.handle is not None:
            lib().par2hip_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now_ms(self) -> float:
        return float(lib().par2hip_server_now_ms(self.handle))

    def submit(self, job: dict, hashes: bool = True, arrival_ms: float = 0.0) -> int:
        e = np.ascontiguousarray(job.get("exps", []), np.uint32)
        miss = np.ascontiguousarray(job.get("missing", []), np.uint32)
        r = Request()
        a = r.job
        a.kind = JOB_REPAIR if job["kind"] == "repair" else JOB_ENCODE
        a.d_slices = job["slices"].ptr
        a.n = job["n"]
        a.slice_size = job["slice_size"]
        a.exps = e.ctypes.data if len(e) else None
        a.count = len(miss) if job["kind"] == "repair" else len(e)
        a.d_recovery = job["recovery"].ptr if job["kind"] == "repair" and job.get("recovery") is not None else None
        a.missing = miss.ctypes.data if len(miss) else None
        r.hashes = 1 if hashes else 0
        r.arrival_ms = arrival_ms
        t = C.c_uint64(0)
        _status(lib().par2hip_server_submit(self.handle, C.byref(r), C.byref(t)), "server_submit")
        self._keep[t.value] = (job,)  # the input buffers must outlive the request
        return t.value

    def wait(self, ticket: int, timeout_ms: float | None = None) -> dict:
        """Wait for a request; with timeout_ms, raise TimeoutError when it is
        not done in time (the request keeps its place)."""
        res = Result()
        if timeout_ms is None:
            rc = lib().par2hip_server_wait(self.handle, ticket, C.byref(res))
        else:
            rc = lib().par2hip_server_wait_timeout(self.handle, ticket, float(timeout_ms), C.byref(res))
            if rc == ETIMEDOUT:
                raise TimeoutError(f"request {ticket} not done within {timeout_ms} ms")
        _status(rc, "server_wait")
        out = {k: getattr(res, k) for k, _ in Result._fields_}
        _status(res.status, f"request {ticket}")
        return out

    def release(self, ticket: int) -> None:
        _status(lib().par2hip_server_release(self.handle, ticket), "server_release")
        self._keep.pop(ticket, None)

    def stats(self) -> dict:
        st = ServerStats()
        _status(lib().par2hip_server_get_stats(self.handle, C.byref(st)), "server_stats")
        return {k: int(getattr(st, k)) for k, _ in ServerStats._fields_}


def d2h(ptr: int, nbytes: int, ctx: "Context") -> np.ndarray:
    """Copy nbytes from a device address (e.g. a server output) to the host."""
    out = np.empty(nbytes, np.uint8)
    ctx._check(lib().par2hip_d2h(ctx.handle, _p(out), C.c_void_p(ptr), nbytes))
    return out


class HipStrategy:
    """GaloisSimdStrategy implemented on the GPU (src/galois/simd/mod.rs:24-128).

    Arrays are numpy uint16 (the reference's ``&mut [u16]`` slices) and are
    updated in place, as the trait does.
    """

    def __init__(self, ctx: Context | None = None):
        self._ctx = ctx

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = Context(0)
        return self._ctx

    def name(self) -> str:
        return lib().par2hip_strategy_name().decode()

    def is_available(self) -> bool:
        return available() > 0

    def priority(self) -> int:
        return int(lib().par2hip_strategy_priority())

    def mul_slice(self, scalar: int, data: np.ndarray) -> None:
        assert data.dtype == np.uint16 and data.flags.c_contiguous
        self.ctx._check(lib().par2hip_gf_mul_slice(self.ctx.handle, scalar, _p(data), data.size))

    def muladd(self, dst: np.ndarray, src: np.ndarray, scalar: int) -> None:
        assert dst.dtype == np.uint16 and src.dtype == np.uint16
        self.ctx._check(lib().par2hip_gf_muladd(self.ctx.handle, _p(dst), _p(src), min(dst.size, src.size), scalar))

    def muladd_region(self, destinations, sources, coefficients, region_offset: int, region_size: int) -> None:
        nd, ns = len(destinations), len(sources)
        if nd == 0:
            return
        dp = (C.c_void_p * nd)(*[d.ctypes.data for d in destinations])
        sp = (C.c_void_p * max(ns, 1))(*[s.ctypes.data for s in sources])
        rows = [np.ascontiguousarray(c, np.uint16) for c in coefficients[:nd]]
        cp = (C.c_void_p * nd)(*[r.ctypes.data if r.size else None for r in rows])
        lens = (C.c_size_t * nd)(*[r.size for r in rows])
        self.ctx._check(lib().par2hip_gf_muladd_region(self.ctx.handle, dp, nd, sp, ns, cp, lens, region_offset,
                                                       region_size))

    # shuffle2x hooks: the GPU consumes interleaved words, so these are the
    # trait's default no-op conversions and plain muladd (simd/mod.rs:67-128)
    def supports_shuffle2x(self) -> bool:
        return False

    def prepare_shuffle2x(self, data) -> None:
        return None

    def finish_shuffle2x(self, data) -> None:
        return None

    def muladd_shuffle2x(self, dst, src, scalar) -> None:
        self.muladd(dst, src, scalar)

    def muladd_region_shuffle2x(self, destinations, sources, coefficients, region_offset, region_size) -> None:
        self.muladd_region(destinations, sources, coefficients, region_offset, region_size)
