"""Python binding of libcolo.so (the C ABI in include/colo.h).

Mirrors the reference's device entry points (Codes/DeviceProcess.h:4-36) and
its driver (colocationFinder::Begin, Codes/colocationFinder.cpp:1068-1147) for
tests and benchmarks.  Device buffers are torch tensors on cuda:0 (memory and
stream plumbing only); all compute runs in the HIP kernels of libcolo.so.
There is no CPU fallback: every call fails loudly when the library or the GPU
is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcolo.so")

COLO_SCHED_PLAIN = 0
COLO_SCHED_PERSISTENT = 1
COLO_SCHED_LEARNED = 2
SCAN_NOALLOC_MAX_N = 1 << 28  # COLO_SCAN_NOALLOC_MAX_N

ERRORS = {0: "ok", 1: "file", 2: "parse", 3: "bounds", 4: "nomem", 5: "hip", 6: "arg", 7: "unsupported"}


class ColoError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"colo error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Dataset(C.Structure):
    _fields_ = [("n", C.c_uint32), ("F", C.c_uint32), ("x", C.c_void_p), ("y", C.c_void_p),
                ("fcount", C.c_void_p), ("minX", C.c_float), ("minY", C.c_float), ("maxX", C.c_float),
                ("maxY", C.c_float), ("labels", C.c_void_p)]


class Grid(C.Structure):
    _fields_ = [("minX", C.c_float), ("minY", C.c_float), ("cellWidth", C.c_float), ("r2", C.c_float),
                ("gx", C.c_uint32), ("gy", C.c_uint32), ("cells", C.c_uint32), ("n", C.c_uint32),
                ("F", C.c_uint32)]


class MineOpts(C.Structure):
    _fields_ = [("theta", C.c_float), ("max_k", C.c_uint32), ("pi_mode", C.c_int32), ("filter_on", C.c_int32),
                ("sched", C.c_int32), ("hash_tables", C.c_int32), ("out_prefix", C.c_char_p),
                ("model", C.c_void_p), ("log_path", C.c_char_p), ("subset_bound", C.c_int32)]


TUNE_UNSET = 0xFFFFFFFF


class Tune(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("count_var", "write_var", "blocks_per_cu", "item_rows")] + \
               [("count_var_k", C.c_uint32 * 16), ("item_rows_k", C.c_uint32 * 16), ("slots_budget_kb", C.c_uint32),
                ("write_rows_k", C.c_uint32 * 16)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        for n in ("count_var_k", "item_rows_k", "write_rows_k"):
            d[n] = list(getattr(self, n))
        return d


class MineStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("filter_ms", "count_ms", "pi_ms", "scan_ms", "write_ms", "host_ms",
                                          "total_ms", "l2_join_ms")] + \
               [(n, C.c_uint64) for n in ("l2_rows", "l2_pairs", "levels", "patterns", "candidates", "launches")] + \
               [(n, C.c_double) for n in ("count_kernel_ms", "write_kernel_ms", "filter_kernel_ms", "pi_kernel_ms",
                                          "count_bytes", "write_bytes")] + \
               [(n, C.c_uint64) for n in ("count_launches", "write_launches")] + \
               [(n, C.c_double) for n in ("l2_count_kernel_ms", "l2_count_bytes", "l2_write_kernel_ms",
                                          "l2_write_bytes", "model_abs_err", "model_n", "arena_bytes")] + \
               [(n, C.c_uint64) for n in ("overflows", "arenas")] + \
               [(n, C.c_double) for n in ("model_ms", "alloc_ms", "alloc_bytes")] + [("level_ms", C.c_double * 16)] + \
               [("level_count_kernel_ms", C.c_double * 16), ("pool_peak_bytes", C.c_double),
                ("slots_deferred", C.c_uint64), ("recounts", C.c_uint64), ("pruned", C.c_uint64),
                ("bound_ms", C.c_double), ("scan_kernel_ms", C.c_double), ("graphs", C.c_uint64),
                ("tables_skipped", C.c_uint64)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["level_ms"] = list(self.level_ms)
        d["level_count_kernel_ms"] = list(self.level_count_kernel_ms)
        return d


_lib = None

SYMBOLS = [
    "colo_ctx_create", "colo_ctx_destroy", "colo_trim", "colo_set_verbose", "colo_get_tune", "colo_set_tune", "colo_autotune", "colo_set_graphs", "colo_graph_info", "colo_bin", "colo_get_grid", "colo_last_error", "colo_binning_export",
    "colo_filter", "colo_bits_words", "colo_join2_count", "colo_scan_slots", "colo_scan_slots_noalloc",
    "colo_join2_write",
    "colo_joink_count", "colo_joink_write", "colo_joink_count_nbr", "colo_joink_write_nbr", "colo_pi", "colo_mine_sweep", "colo_participation", "colo_mine", "colo_free_str",
    "colo_run_config", "colo_load_csv", "colo_load_csv_opts", "colo_host_dataset", "colo_host_label", "colo_host_free",
    "colo_candidates", "colo_mlp_create", "colo_mlp_destroy", "colo_mlp_set_params", "colo_mlp_get_params",
    "colo_mlp_set_norm", "colo_mlp_predict", "colo_mlp_grad", "colo_mlp_train", "colo_mlp_save", "colo_mlp_load",
]


def lib():
    """Loads libcolo.so; raises when it has not been built.

    torch is imported first: torch ships its own HIP runtime under the same
    soname (libamdhip64.so.7) as the one libcolo.so links, and the first one
    loaded serves the whole process.  Loading torch's first keeps torch and
    the kernels of libcolo.so on one runtime, whatever the caller touches
    first (a library loaded before torch put torch on the other runtime and
    made the library's first hipMalloc fail)."""
    global _lib
    if _lib is None:
        import torch  # noqa: F401  (see above)
        if not os.path.exists(LIB_PATH):
            raise ColoError(5, f"{LIB_PATH} is missing: build it first (python gpu-colocationmininig_amd/build_lib.py)")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, f32, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_int
        L.colo_ctx_create.argtypes = [C.POINTER(Dataset), f32, vp, C.POINTER(vp)]
        L.colo_ctx_destroy.argtypes = [vp]
        L.colo_trim.argtypes = [vp]
        L.colo_set_graphs.argtypes = [vp, C.c_int]
        L.colo_set_verbose.argtypes = [vp, C.c_int]
        L.colo_graph_info.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.colo_get_tune.argtypes = [vp, C.POINTER(Tune)]
        L.colo_set_tune.argtypes = [vp, C.POINTER(Tune)]
        L.colo_autotune.argtypes = [vp, C.POINTER(MineOpts), C.c_char_p, C.POINTER(Tune)]
        L.colo_bin.argtypes = [vp, f32, vp]
        L.colo_get_grid.argtypes = [vp, C.POINTER(Grid)]
        L.colo_last_error.restype = C.c_char_p
        L.colo_last_error.argtypes = [vp]
        L.colo_binning_export.argtypes = [vp, vp, vp, vp, C.POINTER(u64), vp]
        L.colo_filter.argtypes = [vp, vp, u32, u32, f32, vp, vp, vp]
        L.colo_bits_words.restype = u64
        L.colo_bits_words.argtypes = [vp, u32, vp]
        L.colo_join2_count.argtypes = [vp, u32, u32, vp, vp, C.POINTER(u64), vp]
        L.colo_scan_slots.argtypes = [vp, vp, u64, vp, C.POINTER(u64), vp]
        L.colo_scan_slots_noalloc.argtypes = [vp, vp, u64, vp, C.POINTER(u64), vp]
        L.colo_join2_write.argtypes = [vp, u32, u32, vp, vp, vp, vp, vp]
        L.colo_joink_count.argtypes = [vp, vp, u64, u32, vp, vp, vp, C.POINTER(u64), vp]
        L.colo_joink_write.argtypes = [vp, vp, u64, u32, vp, vp, vp, vp, vp]
        L.colo_joink_count_nbr.argtypes = [vp, vp, u64, u32, vp, vp, vp, vp, vp, C.POINTER(u64), vp]
        L.colo_joink_write_nbr.argtypes = [vp, vp, u64, u32, vp, vp, vp, vp, vp, vp, vp]
        L.colo_pi.argtypes = [vp, vp, u32, vp, vp, vp, C.POINTER(f32), C.POINTER(f32), vp]
        L.colo_participation.argtypes = [vp, vp, u32, vp, vp, vp]
        L.colo_mine.argtypes = [vp, C.POINTER(MineOpts), C.POINTER(vp), C.POINTER(MineStats)]
        L.colo_mine_sweep.argtypes = [vp, C.POINTER(MineOpts), vp, u32, vp, vp]
        L.colo_free_str.argtypes = [vp]
        L.colo_run_config.argtypes = [C.c_char_p, C.POINTER(MineOpts), C.POINTER(vp)]
        L.colo_load_csv.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.colo_load_csv_opts.argtypes = [C.c_char_p, C.c_char_p, u32, C.POINTER(vp)]
        L.colo_host_dataset.argtypes = [vp, C.POINTER(Dataset)]
        L.colo_host_label.restype = C.c_char_p
        L.colo_host_label.argtypes = [vp, u32]
        L.colo_host_free.argtypes = [vp]
        L.colo_candidates.restype = u64
        L.colo_candidates.argtypes = [vp, u64, u32, vp, u64]
        L.colo_mlp_create.argtypes = [u32, C.POINTER(vp)]
        L.colo_mlp_destroy.argtypes = [vp]
        L.colo_mlp_set_params.argtypes = [vp, vp, C.c_size_t]
        L.colo_mlp_get_params.argtypes = [vp, vp, C.c_size_t]
        L.colo_mlp_set_norm.argtypes = [vp, vp, vp]
        L.colo_mlp_predict.argtypes = [vp, vp, u32, vp]
        L.colo_mlp_grad.argtypes = [vp, vp, vp, u32, vp, C.POINTER(f32)]
        L.colo_mlp_train.argtypes = [vp, vp, vp, u32, u32, u32, f32, u32, C.POINTER(f32)]
        L.colo_mlp_save.argtypes = [vp, C.c_char_p]
        L.colo_mlp_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        for name in SYMBOLS:
            if name in ("colo_free_str", "colo_host_free"):
                getattr(L, name).restype = None
            elif name not in ("colo_last_error", "colo_bits_words", "colo_host_label", "colo_candidates"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise ColoError(5, "no GPU visible: the colocation kernels run on the device only")
    return torch


def _dptr(t):
    return C.c_void_p(t.data_ptr())


def _u32_host(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _on_stream(fn):
    """Runs a Context method with the context's torch stream current, so that
    torch allocations, fills and copies are ordered with the HIP launches."""
    import functools

    @functools.wraps(fn)
    def wrapped(self, *a, **kw):
        with self.torch.cuda.stream(self.tstream):
            return fn(self, *a, **kw)
    return wrapped


class HostData:
    """A parsed dataset CSV (colocationFinder::populateData, cpp:139-254)."""

    def __init__(self, path, cache=None, threads=None):
        """threads: host threads of the parse (None: one per MiB up to the
        hardware's); cache: binary cache file reused while the CSV is
        unchanged."""
        h = C.c_void_p()
        if cache is None and threads is None:
            rc = lib().colo_load_csv(path.encode(), C.byref(h))
        else:
            rc = lib().colo_load_csv_opts(path.encode(), cache.encode() if cache else None, int(threads or 0),
                                          C.byref(h))
        if rc:
            raise ColoError(rc, f"cannot load {path}")
        self.h = h
        ds = Dataset()
        lib().colo_host_dataset(h, C.byref(ds))
        self.n, self.F = ds.n, ds.F
        self.x = np.ctypeslib.as_array(C.cast(ds.x, C.POINTER(C.c_float)), (ds.n,)).copy() if ds.n else np.zeros(0, np.float32)
        self.y = np.ctypeslib.as_array(C.cast(ds.y, C.POINTER(C.c_float)), (ds.n,)).copy() if ds.n else np.zeros(0, np.float32)
        self.fcount = np.ctypeslib.as_array(C.cast(ds.fcount, C.POINTER(C.c_uint32)), (ds.F,)).copy()
        self.bounds = np.array([ds.minX, ds.minY, ds.maxX, ds.maxY], np.float32)
        self.labels = [lib().colo_host_label(h, f).decode() for f in range(ds.F)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().colo_host_free(self.h)
            self.h = None


class Context:
    """Device-resident dataset + grid + binning (the reference's __constant__
    state, Codes/Constant.h:5-22, as an explicit handle)."""

    def __init__(self, x, y, fcount, bounds, r, labels=None):
        torch = _torch()
        self.torch = torch
        L = lib()
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        fcount = _u32_host(fcount)
        bounds = np.asarray(bounds, np.float32)
        self.fcount = fcount
        self.fstart = np.concatenate([[0], np.cumsum(fcount)[:-1]]).astype(np.uint32) if len(fcount) else fcount
        self.n, self.F = len(x), len(fcount)
        ds = Dataset(self.n, self.F, x.ctypes.data, y.ctypes.data, fcount.ctypes.data,
                     float(bounds[0]), float(bounds[1]), float(bounds[2]), float(bounds[3]), None)
        self._labels = None
        if labels is not None:
            arr = (C.c_char_p * self.F)(*[str(l).encode() for l in labels])
            self._labels = arr
            ds.labels = C.cast(arr, C.c_void_p)
        self.tstream = torch.cuda.Stream()
        self.stream = self.tstream.cuda_stream
        h = C.c_void_p()
        rc = L.colo_ctx_create(C.byref(ds), C.c_float(r), C.c_void_p(self.stream), C.byref(h))
        self.h = h
        if rc:
            msg = L.colo_last_error(h).decode() if h else ""
            L.colo_ctx_destroy(h)
            self.h = None
            raise ColoError(rc, msg)
        self.r = r

    @classmethod
    def from_host(cls, hd, r):
        return cls(hd.x, hd.y, hd.fcount, hd.bounds, r, labels=hd.labels)

    def close(self):
        if getattr(self, "h", None):
            lib().colo_ctx_destroy(self.h)
            self.h = None

    def set_verbose(self, level):
        """level > 0: progress lines on stderr (autotuner measurements)."""
        self._check(lib().colo_set_verbose(self.h, int(level)))

    def set_graphs(self, on):
        """Binning replayed from a captured HIP graph and each level's passes
        launched as captured graphs (on), or everything launched eagerly."""
        self._check(lib().colo_set_graphs(self.h, int(bool(on))))

    def graph_info(self):
        """(captures, replays) of the binning graph so far."""
        a, b = C.c_uint32(), C.c_uint32()
        self._check(lib().colo_graph_info(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    @property
    def tune(self):
        """Launch shapes of the batched passes (colo_get_tune) as a dict."""
        t = Tune()
        self._check(lib().colo_get_tune(self.h, C.byref(t)))
        return t.as_dict()

    @tune.setter
    def tune(self, d):
        """Sets launch shapes; a count_var without count_var_k (an
        item_rows without item_rows_k) applies to every level."""
        cur = self.tune
        cur.update(d)
        for one, per in (("count_var", "count_var_k"), ("item_rows", "item_rows_k")):
            if one in d and per not in d:
                cur[per] = [TUNE_UNSET] * 16
        t = Tune(**{k: v for k, v in cur.items() if not k.endswith("_k")})
        for i in range(16):
            t.count_var_k[i] = cur["count_var_k"][i]
            t.item_rows_k[i] = cur["item_rows_k"][i]
            t.write_rows_k[i] = cur["write_rows_k"][i]
        self._check(lib().colo_set_tune(self.h, C.byref(t)))

    @_on_stream
    def autotune(self, theta, max_k=0, sched=COLO_SCHED_PERSISTENT, model=None, cache=None, pi_mode=0,
                 filter_on=True, subset_bound=False):
        """Chooses the launch shapes on this context's data (colo_autotune)."""
        o = _opts(theta, max_k, pi_mode, filter_on, sched, False, None, model, None, subset_bound)
        t = Tune()
        self._check(lib().colo_autotune(self.h, C.byref(o), cache.encode() if cache else None, C.byref(t)))
        return t.as_dict()

    def trim(self):
        """Returns the device memory cached by earlier mining runs (colo_trim)."""
        self._check(lib().colo_trim(self.h))

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise ColoError(rc, lib().colo_last_error(self.h).decode())

    def _u32(self, n):
        return self.torch.zeros(max(int(n), 1), dtype=self.torch.int32, device="cuda")

    @_on_stream
    def _np_u32(self, t, n=None):
        """Device tensor -> host uint32 array, ordered after the context's
        launches (the copy runs on the context stream)."""
        a = t.cpu().numpy().view(np.uint32)
        return a if n is None else a[:n]

    @_on_stream
    def rebin(self, r):
        """setGrid + CountCellFeatures on the resident data (cpp:102-126)."""
        self._check(lib().colo_bin(self.h, C.c_float(r), C.c_void_p(self.stream)))
        self.r = r

    @_on_stream
    def grid(self):
        g = Grid()
        self._check(lib().colo_get_grid(self.h, C.byref(g)))
        return {n: getattr(g, n) for n, _ in Grid._fields_}

    @_on_stream
    def binning(self):
        """cell of every instance (input order), instances in sorted order,
        dense directory F x (cells+1)."""
        dl = C.c_uint64()
        self._check(lib().colo_binning_export(self.h, None, None, None, C.byref(dl), C.c_void_p(self.stream)))
        cell, sidx, d = self._u32(self.n), self._u32(self.n), self._u32(dl.value)
        self._check(lib().colo_binning_export(self.h, _dptr(cell), _dptr(sidx), _dptr(d), C.byref(dl),
                                              C.c_void_p(self.stream)))
        return self._np_u32(cell, self.n), self._np_u32(sidx, self.n), self._np_u32(d, dl.value)

    @_on_stream
    def filter(self, cands, k, theta):
        cands = _u32_host(cands).reshape(-1)
        Cn = len(cands) // k
        torch = self.torch
        dc = torch.from_numpy(cands.view(np.int32)).cuda() if Cn else self._u32(1)
        counts = self._u32(Cn * k)
        passed = torch.zeros(max(Cn, 1), dtype=torch.uint8, device="cuda")
        self._check(lib().colo_filter(self.h, _dptr(dc), Cn, k, C.c_float(theta), _dptr(counts), _dptr(passed),
                                      C.c_void_p(self.stream)))
        return self._np_u32(counts, Cn * k).reshape(Cn, k), passed.cpu().numpy()[:Cn].astype(bool)

    def bits_words(self, cand):
        cand = _u32_host(cand)
        return lib().colo_bits_words(self.h, len(cand), cand.ctypes.data)

    @_on_stream
    def join2_count(self, a, b):
        Na = int(self.fcount[a])
        slots = self._u32(Na)
        bits = self._u32(self.bits_words([a, b]))
        rows = C.c_uint64()
        self._check(lib().colo_join2_count(self.h, a, b, _dptr(slots), _dptr(bits), C.byref(rows),
                                           C.c_void_p(self.stream)))
        return rows.value, slots, bits

    @_on_stream
    def scan(self, slots, n):
        """Exclusive scan of n slot counts: n + 1 offsets (the sum last)."""
        offs = self.torch.zeros(int(n) + 1, dtype=self.torch.int64, device="cuda")
        tot = C.c_uint64()
        self._check(lib().colo_scan_slots(self.h, _dptr(slots), n, _dptr(offs), C.byref(tot), C.c_void_p(self.stream)))
        return offs, tot.value

    @_on_stream
    def join2_write(self, a, b, slots, lists=False):
        """Level-2 table (and, with lists=True, the neighbour lists as sorted
        positions plus the n+1 row offsets)."""
        Na = int(self.fcount[a])
        offs, tot = self.scan(slots, Na)
        table = self._u32(2 * tot)
        nbr = self._u32(tot) if lists else None
        self._check(lib().colo_join2_write(self.h, a, b, _dptr(slots), _dptr(offs), _dptr(table),
                                           _dptr(nbr) if lists else None, C.c_void_p(self.stream)))
        if lists:
            return table, tot, nbr, offs
        return table, tot

    @_on_stream
    def joink_count(self, parent, R, k, cand):
        cand = _u32_host(cand)
        slots = self._u32(R)
        bits = self._u32(self.bits_words(cand))
        rows = C.c_uint64()
        self._check(lib().colo_joink_count(self.h, _dptr(parent), R, k, cand.ctypes.data, _dptr(slots), _dptr(bits),
                                           C.byref(rows), C.c_void_p(self.stream)))
        return rows.value, slots, bits

    @_on_stream
    def joink_write(self, parent, R, k, cand, slots):
        cand = _u32_host(cand)
        offs, tot = self.scan(slots, R)
        table = self._u32(k * tot)
        self._check(lib().colo_joink_write(self.h, _dptr(parent), R, k, cand.ctypes.data, _dptr(slots), _dptr(offs),
                                           _dptr(table), C.c_void_p(self.stream)))
        return table, tot

    @_on_stream
    def joink_count_nbr(self, parent, R, k, cand, table2, offs2):
        """Level-k count with candidates from the anchors' neighbour lists
        (table2 / offs2: neighbour lists of (cand[0], cand[-1]) from
        join2_write(lists=True) and their n+1 row offsets)."""
        cand = _u32_host(cand)
        slots = self._u32(R)
        bits = self._u32(self.bits_words(cand))
        rows = C.c_uint64()
        self._check(lib().colo_joink_count_nbr(self.h, _dptr(parent), R, k, cand.ctypes.data, _dptr(table2),
                                               _dptr(offs2), _dptr(slots), _dptr(bits), C.byref(rows),
                                               C.c_void_p(self.stream)))
        return rows.value, slots, bits

    @_on_stream
    def joink_write_nbr(self, parent, R, k, cand, slots, table2, offs2):
        cand = _u32_host(cand)
        offs, tot = self.scan(slots, R)
        table = self._u32(k * tot)
        self._check(lib().colo_joink_write_nbr(self.h, _dptr(parent), R, k, cand.ctypes.data, _dptr(table2),
                                               _dptr(offs2), _dptr(slots), _dptr(offs), _dptr(table),
                                               C.c_void_p(self.stream)))
        return table, tot

    @_on_stream
    def pi(self, bits, cand):
        cand = _u32_host(cand)
        k = len(cand)
        nc = np.zeros(k, np.uint64)
        ne = np.zeros(k, np.uint64)
        pc, pe = C.c_float(), C.c_float()
        self._check(lib().colo_pi(self.h, _dptr(bits), k, cand.ctypes.data, nc.ctypes.data, ne.ctypes.data,
                                  C.byref(pc), C.byref(pe), C.c_void_p(self.stream)))
        return nc, ne, pc.value, pe.value

    @_on_stream
    def participation(self, bits, cand):
        cand = _u32_host(cand)
        tot = int(sum(int(self.fcount[f]) for f in cand))
        flags = self.torch.zeros(max(tot, 1), dtype=self.torch.uint8, device="cuda")
        self._check(lib().colo_participation(self.h, _dptr(bits), len(cand), cand.ctypes.data, _dptr(flags),
                                             C.c_void_p(self.stream)))
        return flags.cpu().numpy()[:tot]

    @_on_stream
    def mine(self, theta, max_k=0, pi_mode=0, filter_on=True, sched=COLO_SCHED_PERSISTENT, hash_tables=True,
             out_prefix=None, model=None, log_path=None, subset_bound=False):
        o = _opts(theta, max_k, pi_mode, filter_on, sched, hash_tables, out_prefix, model, log_path, subset_bound)
        rep = C.c_void_p()
        st = MineStats()
        rc = lib().colo_mine(self.h, C.byref(o), C.byref(rep), C.byref(st))
        text = None
        if rep.value:
            text = C.string_at(rep.value).decode()
            lib().colo_free_str(rep)
        self._check(rc)
        return text, st.as_dict()

    @_on_stream
    def mine_sweep(self, thetas, max_k=0, pi_mode=0, filter_on=True, sched=COLO_SCHED_PERSISTENT, hash_tables=True,
                   out_prefix=None, model=None, subset_bound=False):
        """Mines every theta in one call, reusing theta-independent results
        (colo_mine_sweep); returns [(report, stats)] in the given order."""
        o = _opts(0.0, max_k, pi_mode, filter_on, sched, hash_tables, out_prefix, model, None, subset_bound)
        n = len(thetas)
        th = np.ascontiguousarray(thetas, np.float32)
        reps = (C.c_void_p * n)()
        sts = (MineStats * n)()
        rc = lib().colo_mine_sweep(self.h, C.byref(o), th.ctypes.data, n, C.cast(reps, C.c_void_p),
                                   C.cast(sts, C.c_void_p))
        out = []
        for i in range(n):
            text = None
            if reps[i]:
                text = C.string_at(reps[i]).decode()
                lib().colo_free_str(reps[i])
            out.append((text, sts[i].as_dict()))
        try:
            self._check(rc)
        except ColoError as e:
            e.results = out
            raise
        return out


MLP_IN, MLP_HIDDEN, MLP_OUT = 16, 64, 2
MLP_NPARAMS = 64 * 16 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2
FEATURE_NAMES = ["k", "log1p N_first", "log1p N_last", "log1p N_min", "log1p N_max", "log1p anchors",
                 "log1p filter_first", "log1p filter_last", "filter bound", "log1p cells_first", "log1p cells_last",
                 "log1p 3x3 work(first,last)", "log1p R2(first,last)", "log1p R2(prev,last)",
                 "log occupancy_first", "log occupancy_last"]
OUTPUT_NAMES = ["log1p rows", "log1p refinement tests"]
RECORD_FLOATS = MLP_IN + MLP_OUT + 1  # features, log1p(rows), log1p(tests), k


def _opts(theta, max_k, pi_mode, filter_on, sched, hash_tables, out_prefix, model, log_path, subset_bound=False):
    return MineOpts(C.c_float(theta), max_k, pi_mode, int(filter_on), sched, int(hash_tables),
                    out_prefix.encode() if out_prefix else None, model.h if model is not None else None,
                    log_path.encode() if log_path else None, int(subset_bound))


def read_log(path, with_k=False):
    """Training records written by the miner's log_path hook: features
    (n x MLP_IN), targets (n x MLP_OUT: log1p rows, log1p refinement tests)
    and, with with_k, the pattern size of each record."""
    a = np.fromfile(path, dtype=np.float32)
    a = a[: len(a) // RECORD_FLOATS * RECORD_FLOATS].reshape(-1, RECORD_FLOATS)
    X, T = a[:, :MLP_IN].copy(), a[:, MLP_IN:MLP_IN + MLP_OUT].copy()
    return (X, T, a[:, -1].astype(np.int64)) if with_k else (X, T)


class CostModel:
    """FP32 MLP 16 -> 64 -> 64 -> 64 -> 2 on the GPU (HIP forward, backward
    and Adam kernels in colo_mlp.hip) predicting log1p(table rows) and
    log1p(refinement tests) of a candidate."""

    def __init__(self, seed=0, h=None):
        _torch()
        if h is None:
            h = C.c_void_p()
            rc = lib().colo_mlp_create(seed, C.byref(h))
            if rc:
                raise ColoError(rc, "colo_mlp_create")
        self.h = h

    @classmethod
    def load(cls, path):
        _torch()
        h = C.c_void_p()
        rc = lib().colo_mlp_load(path.encode(), C.byref(h))
        if rc:
            raise ColoError(rc, f"cannot load {path}")
        return cls(h=h)

    def save(self, path):
        rc = lib().colo_mlp_save(self.h, path.encode())
        if rc:
            raise ColoError(rc, f"cannot save {path}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().colo_mlp_destroy(self.h)
            self.h = None

    @property
    def params(self):
        p = np.zeros(MLP_NPARAMS, np.float32)
        rc = lib().colo_mlp_get_params(self.h, p.ctypes.data, MLP_NPARAMS)
        if rc:
            raise ColoError(rc, "get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        rc = lib().colo_mlp_set_params(self.h, p.ctypes.data, len(p))
        if rc:
            raise ColoError(rc, "set_params")

    def set_norm(self, mean, std):
        mean = np.ascontiguousarray(mean, np.float32)
        std = np.ascontiguousarray(std, np.float32)
        rc = lib().colo_mlp_set_norm(self.h, mean.ctypes.data, std.ctypes.data)
        if rc:
            raise ColoError(rc, "set_norm")

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros((len(X), MLP_OUT), np.float32)
        rc = lib().colo_mlp_predict(self.h, X.ctypes.data, len(X), y.ctypes.data)
        if rc:
            raise ColoError(rc, "predict")
        return y

    def grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32).reshape(len(X), MLP_OUT)
        g = np.zeros(MLP_NPARAMS, np.float32)
        loss = C.c_float()
        rc = lib().colo_mlp_grad(self.h, X.ctypes.data, T.ctypes.data, len(X), g.ctypes.data, C.byref(loss))
        if rc:
            raise ColoError(rc, "grad")
        return g, loss.value

    def train(self, X, T, steps=2000, batch=256, lr=1e-3, seed=0, normalise=True):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32).reshape(len(X), MLP_OUT)
        if normalise:
            self.set_norm(X.mean(0), X.std(0) + 1e-6)
        loss = C.c_float()
        rc = lib().colo_mlp_train(self.h, X.ctypes.data, T.ctypes.data, len(X), steps, batch, C.c_float(lr), seed,
                                  C.byref(loss))
        if rc:
            raise ColoError(rc, "train")
        return loss.value


def candidates(L, km1):
    """candidateColocationGeneral (cpp:517-557) on the host: returns the
    next-level candidates of the prevalent list L (rows of km1 ids)."""
    L = _u32_host(L).reshape(-1)
    P = len(L) // km1 if km1 else 0
    n = lib().colo_candidates(L.ctypes.data, P, km1, None, 0)
    out = np.zeros(max(n * (km1 + 1), 1), np.uint32)
    lib().colo_candidates(L.ctypes.data, P, km1, out.ctypes.data, n)
    return out[: n * (km1 + 1)].reshape(n, km1 + 1)


def run_config(path, sched=COLO_SCHED_PERSISTENT, max_k=0, pi_mode=0, filter_on=True):
    """colocationFinder::Begin(argc, argv): config.txt -> pattern files."""
    _torch()
    o = MineOpts(0.0, max_k, pi_mode, int(filter_on), sched, 0, None)
    line = C.c_void_p()
    rc = lib().colo_run_config(path.encode(), C.byref(o), C.byref(line))
    if rc:
        raise ColoError(rc, f"run_config({path})")
    s = C.string_at(line.value).decode()
    lib().colo_free_str(line)
    return s
