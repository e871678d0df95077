"""Shared test inputs: the QuickPar golden set and seeded synthetic slices."""
from __future__ import annotations

import os

import numpy as np

import par2fmt

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
QUICKPAR = os.path.join(GOLDEN, "quickpar")


class QuickPar:
    """The reference's fixture as slices in canonical (Main packet) order."""

    def __init__(self):
        self.set = par2fmt.load_dir(QUICKPAR)
        self.slice_size = self.set.slice_size
        files = self.set.ordered_files()
        slices, ifsc, owner = [], [], []
        for fi, f in enumerate(files):
            with open(os.path.join(QUICKPAR, f.name), "rb") as fh:
                b = fh.read()
            assert len(b) == f.length
            nsl = -(-len(b) // self.slice_size)
            for k in range(nsl):
                chunk = b[k * self.slice_size:(k + 1) * self.slice_size]
                slices.append(chunk + bytes(self.slice_size - len(chunk)))
                owner.append((fi, k))
            ifsc.extend(self.set.ifsc[f.file_id])
        self.files = files
        self.owner = owner
        self.data = np.frombuffer(b"".join(slices), np.uint8).reshape(len(slices), self.slice_size).copy()
        self.ifsc = ifsc
        self.exponents = sorted(self.set.recovery)
        self.recovery = np.stack([np.frombuffer(self.set.recovery[e], np.uint8) for e in self.exponents])

    @property
    def words(self):
        return self.data.view(np.uint16)

    @property
    def rec_words(self):
        return self.recovery.view(np.uint16)


_qp = None


def quickpar() -> QuickPar:
    global _qp
    if _qp is None:
        _qp = QuickPar()
    return _qp


def xoshiro_bytes(seed: int, nbytes: int) -> np.ndarray:
    """Seeded uniform bytes (numpy PCG64 stream; deterministic across hosts)."""
    return np.random.default_rng(seed).integers(0, 256, nbytes, dtype=np.uint8)
