"""ctypes wrapper of the CPU oracle (oracle/colo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker, never as the
thing measured for the GPU path.  See colo_oracle.c for the reference lines
each function restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_load_csv.restype = vp
        L.orc_load_csv.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.orc_from_arrays.restype = vp
        L.orc_from_arrays.argtypes = [C.c_uint32, C.c_uint32, _f32p, _f32p, _u32p, _f32p]
        L.orc_free.argtypes = [vp]
        L.orc_n.restype = C.c_uint32
        L.orc_n.argtypes = [vp]
        L.orc_F.restype = C.c_uint32
        L.orc_F.argtypes = [vp]
        L.orc_features.argtypes = [vp, _u32p, _u32p]
        L.orc_points.argtypes = [vp, _f32p, _f32p]
        L.orc_bounds.argtypes = [vp, _f32p]
        L.orc_label.restype = C.c_char_p
        L.orc_label.argtypes = [vp, C.c_uint32]
        L.orc_load_params.restype = C.c_int
        L.orc_load_params.argtypes = [C.c_char_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.POINTER(C.c_uint32)]
        L.orc_bin.restype = C.c_int
        L.orc_bin.argtypes = [vp, C.c_float]
        L.orc_grid.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
        L.orc_binning.argtypes = [vp, _u32p, _u32p]
        L.orc_cell_counts.argtypes = [vp, _i32p, _i32p]
        L.orc_nbr_cell.restype = C.c_int32
        L.orc_nbr_cell.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.c_int32]
        L.orc_join2_count.restype = C.c_uint64
        L.orc_join2_count.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp, vp]
        L.orc_join2_write.argtypes = [vp, C.c_uint32, C.c_uint32, _u32p, _u32p]
        L.orc_joink_count.restype = C.c_uint64
        L.orc_joink_count.argtypes = [vp, _u32p, C.c_uint64, C.c_uint32, _u32p, vp, vp]
        L.orc_joink_write.argtypes = [vp, _u32p, C.c_uint64, C.c_uint32, _u32p, _u32p, _u32p]
        L.orc_pi_compat.restype = C.c_float
        L.orc_pi_compat.argtypes = [vp, _u32p, C.c_uint32, _u32p, _u64p]
        L.orc_pi_exact.restype = C.c_float
        L.orc_pi_exact.argtypes = [vp, _u32p, C.c_uint32, _u32p, _u64p]
        L.orc_filter.argtypes = [vp, _u32p, C.c_uint32, C.c_uint32, C.c_float, _u32p, _u8p]
        L.orc_filter_literal.argtypes = [vp, _u32p, C.c_uint32, C.c_uint32, C.c_float, _u32p, _u8p]
        L.orc_candgen.restype = C.c_uint64
        L.orc_candgen.argtypes = [_u32p, C.c_uint64, C.c_uint32, vp, C.c_uint64]
        L.orc_table_hash.restype = C.c_uint64
        L.orc_table_hash.argtypes = [_u32p, C.c_uint64, C.c_uint32]
        L.orc_row_hash.restype = C.c_uint64
        L.orc_row_hash.argtypes = [_u32p, C.c_uint32]
        L.orc_mine.restype = C.c_void_p
        L.orc_mine.argtypes = [vp, C.c_float, C.c_float, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_char_p, C.POINTER(C.c_int)]
        L.orc_free_str.argtypes = [C.c_void_p]
        L.orc_boundary_pairs.argtypes = [vp, _u64p]
        L.orc_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One dataset + grid, as held by colocationFinder after populateData/setGrid."""

    def __init__(self, handle):
        self.h = handle
        L = lib()
        self.n = L.orc_n(handle)
        self.F = L.orc_F(handle)
        self.fstart = np.zeros(self.F, np.uint32)
        self.fcount = np.zeros(self.F, np.uint32)
        if self.F:
            L.orc_features(handle, self.fstart, self.fcount)
        self.r = None

    @classmethod
    def from_csv(cls, path):
        err = C.c_int(0)
        h = lib().orc_load_csv(path.encode(), C.byref(err))
        if not h:
            raise ValueError(f"oracle: cannot parse {path} (code {err.value})")
        return cls(h)

    @classmethod
    def from_arrays(cls, x, y, fcount, bounds):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        fcount = np.ascontiguousarray(fcount, np.uint32)
        bounds = np.ascontiguousarray(bounds, np.float32)
        h = lib().orc_from_arrays(len(x), len(fcount), x, y, fcount, bounds)
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def points(self):
        x = np.zeros(self.n, np.float32)
        y = np.zeros(self.n, np.float32)
        lib().orc_points(self.h, x, y)
        return x, y

    def bounds(self):
        b = np.zeros(4, np.float32)
        lib().orc_bounds(self.h, b)
        return b

    def label(self, f):
        return lib().orc_label(self.h, f).decode()

    def bin(self, r):
        rc = lib().orc_bin(self.h, r)
        if rc:
            raise ValueError(f"oracle: binning failed (code {rc})")
        self.r = r

    def grid(self):
        gx, gy, cells, r2 = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_float()
        lib().orc_grid(self.h, C.byref(gx), C.byref(gy), C.byref(cells), C.byref(r2))
        return gx.value, gy.value, cells.value, r2.value

    def binning(self):
        cid = np.zeros(self.n, np.uint32)
        sidx = np.zeros(self.n, np.uint32)
        lib().orc_binning(self.h, cid, sidx)
        return cid, sidx

    def cell_counts(self):
        _, _, cells, _ = self.grid()
        cnt = np.zeros(cells * self.F, np.int32)
        st = np.zeros(cells * self.F, np.int32)
        lib().orc_cell_counts(self.h, cnt, st)
        return cnt, st

    def join2_count(self, a, b):
        Na, Nb = int(self.fcount[a]), int(self.fcount[b])
        slots = np.zeros(max(Na, 1), np.uint32)
        bm = np.zeros(Na + Nb + 1, np.uint32)
        ex = np.zeros(Na + Nb + 1, np.uint32)
        rows = lib().orc_join2_count(self.h, a, b, _ptr(slots), _ptr(bm), _ptr(ex))
        return rows, slots[:Na], bm, ex

    def join2_write(self, a, b, slots):
        slots = np.ascontiguousarray(slots, np.uint32)
        rows = int(slots.sum(dtype=np.uint64))
        table = np.zeros(max(2 * rows, 1), np.uint32)
        lib().orc_join2_write(self.h, a, b, slots, table)
        return table[: 2 * rows]

    def joink_count(self, parent, k, cand):
        parent = np.ascontiguousarray(parent, np.uint32)
        cand = np.ascontiguousarray(cand, np.uint32)
        R = len(parent) // (k - 1)
        slots = np.zeros(max(R, 1), np.uint32)
        wsz = int(sum(int(self.fcount[f]) for f in cand))
        bm = np.zeros(wsz + 1, np.uint32)
        rows = lib().orc_joink_count(self.h, parent, R, k, cand, _ptr(slots), _ptr(bm))
        return rows, slots[:R], bm

    def joink_write(self, parent, k, cand, slots):
        parent = np.ascontiguousarray(parent, np.uint32)
        cand = np.ascontiguousarray(cand, np.uint32)
        slots = np.ascontiguousarray(slots, np.uint32)
        R = len(parent) // (k - 1)
        rows = int(slots.sum(dtype=np.uint64))
        table = np.zeros(max(k * rows, 1), np.uint32)
        lib().orc_joink_write(self.h, parent, R, k, cand, slots, table)
        return table[: k * rows]

    def pi_compat(self, bitmap, cand):
        cand = np.ascontiguousarray(cand, np.uint32)
        num = np.zeros(len(cand), np.uint64)
        pi = lib().orc_pi_compat(self.h, np.ascontiguousarray(bitmap, np.uint32), len(cand), cand, num)
        return pi, num

    def pi_exact(self, bitmap, cand):
        cand = np.ascontiguousarray(cand, np.uint32)
        num = np.zeros(len(cand), np.uint64)
        pi = lib().orc_pi_exact(self.h, np.ascontiguousarray(bitmap, np.uint32), len(cand), cand, num)
        return pi, num

    def filter(self, cands, k, theta, literal=False):
        cands = np.ascontiguousarray(cands, np.uint32).reshape(-1)
        Cn = len(cands) // k
        counts = np.zeros(max(Cn * k, 1), np.uint32)
        passed = np.zeros(max(Cn, 1), np.uint8)
        fn = lib().orc_filter_literal if literal else lib().orc_filter
        fn(self.h, cands, Cn, k, theta, counts, passed)
        return counts[: Cn * k].reshape(Cn, k), passed[:Cn].astype(bool)

    def boundary_pairs(self):
        """Distance tests of the level-2 scan where a fused or a
        double-accumulated squared distance decides d2 <= r2 differently from
        the canonical binary32 form (orc_boundary_pairs)."""
        out = np.zeros(6, np.uint64)
        lib().orc_boundary_pairs(self.h, out)
        keys = ("tests", "neighbours", "fma_dx_first", "fma_dy_first", "double_sum", "any_form")
        return {k: int(v) for k, v in zip(keys, out)}

    def mine(self, r, theta, max_k=0, pi_mode=0, filter_on=True, hash_tables=True, out_prefix=None,
             subset_bound=False):
        err = C.c_int(0)
        p = lib().orc_mine(self.h, r, theta, max_k, pi_mode, int(filter_on), int(hash_tables), int(subset_bound),
                           out_prefix.encode() if out_prefix else None, C.byref(err))
        if not p:
            raise ValueError(f"oracle: mining failed (code {err.value})")
        s = C.string_at(p).decode()
        lib().orc_free_str(p)
        return s


def nbr_cell(gx, cells, cur, i, j):
    return lib().orc_nbr_cell(gx, cells, cur, i, j)


def candgen(L, km1):
    L = np.ascontiguousarray(L, np.uint32).reshape(-1)
    P = len(L) // km1 if km1 else 0
    n = lib().orc_candgen(L, P, km1, None, 0)
    out = np.zeros(max(n * (km1 + 1), 1), np.uint32)
    lib().orc_candgen(L, P, km1, out.ctypes.data_as(C.c_void_p), n)
    return out[: n * (km1 + 1)].reshape(n, km1 + 1)


def table_hash(t):
    """Row-order independent checksum of a rows x k table (orc_table_hash)."""
    t = np.ascontiguousarray(t, np.uint32)
    if t.ndim != 2:
        raise ValueError("table_hash takes a rows x k table")
    return lib().orc_table_hash(t.reshape(-1), t.shape[0], t.shape[1])


def row_hash(row):
    row = np.ascontiguousarray(row, np.uint32)
    return lib().orc_row_hash(row, len(row))


def load_params(path):
    r, th, fl = C.c_float(), C.c_float(), C.c_uint32()
    rc = lib().orc_load_params(path.encode(), C.byref(r), C.byref(th), C.byref(fl))
    if rc:
        raise ValueError(f"oracle: cannot read {path}")
    return r.value, th.value, fl.value


def threads():
    return lib().orc_threads()


def set_threads(t):
    lib().orc_set_threads(t)
