"""Python binding of the storm_hip C ABI (include/storm_hip.h) over ctypes.

Host-side helper for the tests and the benchmark: it loads the in-tree
libstorm_hip.so built for gfx950 and raises if the library is missing or a
call fails. There is no fallback of any kind.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstorm_hip.so")
GPU_LIB_PATH = os.path.join(HERE, "libstorm_gpu.so")

POLICY = {"all_bitmap": 0, "analytic": 1, "learned": 2, "fixed": 3, "exhaustive": 4,
          "probe": 5, "refine": 6, "per_pair": 7}
OUT = {"sum": 0, "tri": 1, "tri_density": 2}
BB = {"auto": 0, "valu": 1, "mfma": 2, "mfma_fp4": 3}
CELLS = ["BB", "BS", "BR", "BW", "SS", "SR", "SW", "RR", "RW", "WW", "EMPTY", "CB"]
NCELL = len(CELLS)
# Representation flags of from_lists (storm_hip.h STORM_HIP_BUILD_*).
BUILD_RUNS, BUILD_EWAH, BUILD_COMP, BUILD_ON_HOST, BUILD_SLICES = 4, 8, 16, 32, 64
# Arena ids of storm_hip_corpus_download_arena, in order, with element types.
ARENAS = [("list_off", np.uint64), ("list", np.uint32), ("run_off", np.uint64),
          ("run_start", np.uint32), ("run_end", np.uint32), ("ewah_off", np.uint64),
          ("ewah", np.uint64), ("mk_off", np.uint64), ("mk_word", np.uint32),
          ("mk_pos", np.uint32), ("comp_off", np.uint64), ("comp", np.uint32),
          ("meta", np.uint32), ("bm_row", np.uint32), ("bitmap", np.uint64),
          ("rank_row", np.uint32), ("rank", np.uint64), ("occ_row", np.uint32),
          ("occ", np.uint64)]
GEN = {"zipf": 0, "runs": 1, "bernoulli": 2}


class StormHipError(RuntimeError):
    pass


class Options(C.Structure):
    _fields_ = [("policy", C.c_uint32), ("out_mode", C.c_uint32), ("bb_kernel", C.c_uint32),
                ("tile", C.c_uint32), ("fixed_cell", C.c_uint32), ("use_graph", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class Stats(C.Structure):
    _fields_ = [("sum", C.c_uint64), ("pairs", C.c_uint64), ("skipped", C.c_uint64),
                ("decisions", C.c_uint64), ("cell_pairs", C.c_uint64 * NCELL),
                ("cell_tiles", C.c_uint64 * NCELL), ("fallback_tiles", C.c_uint64),
                ("ms_total", C.c_double), ("ms_device", C.c_double), ("ms_bb", C.c_double),
                ("ms_select", C.c_double), ("bb_launches", C.c_uint64),
                ("probe_tiles", C.c_uint64), ("probe_switched", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_
             if k not in ("cell_pairs", "cell_tiles")}
        d["cell_pairs"] = {CELLS[i]: int(self.cell_pairs[i]) for i in range(NCELL)
                           if self.cell_pairs[i]}
        d["cell_tiles"] = {CELLS[i]: int(self.cell_tiles[i]) for i in range(NCELL)
                           if self.cell_tiles[i]}
        return d


class PrefixInfo(C.Structure):
    _fields_ = [("est_ratio", C.c_double), ("prefix_elements", C.c_uint64),
                ("candidates", C.c_uint64), ("unique_candidates", C.c_uint64),
                ("used_prefix", C.c_uint32)]


class Arena(C.Structure):
    P64 = C.POINTER(C.c_uint64)
    P32 = C.POINTER(C.c_uint32)
    _fields_ = [("n_rows", C.c_uint32), ("universe", C.c_uint32),
                ("list_off", P64), ("list", P32),
                ("run_off", P64), ("run_start", P32), ("run_end", P32),
                ("ewah_off", P64), ("ewah", P64),
                ("bitmap_off", P64), ("bitmap", P64),
                ("rank_off", P64), ("rank", P64),
                ("occ_off", P64), ("occ", P64),
                ("comp_off", P64), ("comp", P32),
                ("meta", P32)]


class CostModel(C.Structure):
    _fields_ = [("ns_per_unit", C.c_double * 10), ("ns_fixed", C.c_double),
                ("ns_per_occ_word", C.c_double), ("ns_probe_hot", C.c_double),
                ("ns_probe_cold", C.c_double), ("bb_zonemap", C.c_double)]

    def as_dict(self):
        return {"ns_per_unit": {CELLS[i]: self.ns_per_unit[i] for i in range(10)},
                "ns_fixed": self.ns_fixed, "ns_per_occ_word": self.ns_per_occ_word,
                "bb_zonemap": self.bb_zonemap}


class BBConfig(C.Structure):
    _fields_ = [("stage_words", C.c_uint32), ("band", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("waves", C.c_uint32)]

    def as_dict(self):
        return {"stage_words": self.stage_words, "band": self.band,
                "blocks_per_cu": self.blocks_per_cu, "waves": self.waves}


class TuneResult(C.Structure):
    _fields_ = [("tile", C.c_uint32), ("n_tried", C.c_uint32), ("bb", BBConfig),
                ("ms_best", C.c_double), ("ms_default", C.c_double)]

    def as_dict(self):
        return {"tile": self.tile, "n_tried": self.n_tried, "bb": self.bb.as_dict(),
                "ms_best": self.ms_best, "ms_default": self.ms_default}


_lib = None


def lib():
    """Load libstorm_hip.so (built by `make -C stormbitmaps-internal_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise StormHipError(f"{LIB_PATH} is missing: build it with "
                            "`make -C stormbitmaps-internal_amd` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    pvp = C.POINTER(C.c_void_p)
    sig = {
        "storm_hip_strerror": (C.c_char_p, [i32]),
        "storm_hip_version": (C.c_char_p, []),
        "storm_hip_last_error": (C.c_char_p, []),
        "storm_hip_device_count": (i32, [C.POINTER(i32)]),
        "storm_hip_ctx_create": (i32, [i32, u32, pvp]),
        "storm_hip_ctx_free": (None, [vp]),
        "storm_hip_sync": (i32, [vp]),
        "storm_hip_corpus_upload_contig": (i32, [vp, vp, u32, u32, pvp]),
        "storm_hip_corpus_gen_bernoulli": (i32, [vp, u32, u32, C.c_double, u64, pvp]),
        "storm_hip_corpus_download_words": (i32, [vp, vp, u32, u32, vp]),
        "storm_hip_corpus_upload": (i32, [vp, C.POINTER(Arena), pvp]),
        "storm_hip_corpus_from_lists": (i32, [vp, vp, vp, u32, u32, u32, pvp]),
        "storm_hip_corpus_materialize_bitmaps": (i32, [vp, vp]),
        "storm_hip_corpus_free": (None, [vp]),
        "storm_hip_corpus_rows": (u32, [vp]),
        "storm_hip_corpus_device_bytes": (u64, [vp]),
        "storm_hip_allpairs": (i32, [vp, vp, C.POINTER(Options), C.POINTER(Stats)]),
        "storm_hip_result_copy_tri": (i32, [vp, vp, u64, u64, vp]),
        "storm_hip_result_row_sums": (i32, [vp, vp, vp]),
        "storm_hip_set_cost_model": (i32, [vp, C.POINTER(CostModel)]),
        "storm_hip_get_cost_model": (i32, [vp, C.POINTER(CostModel)]),
        "storm_hip_last_decisions": (i32, [vp, vp, vp, u64]),
        "storm_hip_gen_corpus": (i32, [u32, u32, u32, C.c_double, C.c_double, C.c_double, u64,
                                       C.POINTER(C.POINTER(C.c_uint64)),
                                       C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(u64)]),
        "storm_hip_gen_free": (None, [vp]),
        "storm_hip_mlp_shape": (i32, [C.POINTER(i32)] * 5),
        "storm_hip_selector_sample": (i32, [vp, vp, u32, u32, u64, u32, vp, vp, vp, vp,
                                            C.POINTER(u32)]),
        "storm_hip_selector_train": (i32, [vp, vp, vp, vp, u32, u32, u32, C.c_float, u64,
                                           C.POINTER(C.c_float)]),
        "storm_hip_selector_predict": (i32, [vp, vp, u32, vp]),
        "storm_hip_selector_set_params": (i32, [vp, vp, vp]),
        "storm_hip_selector_get_params": (i32, [vp, vp, vp]),
        "storm_hip_selector_grad": (i32, [vp, vp, vp, vp, u32, vp, C.POINTER(C.c_float)]),
        "storm_hip_selector_adam_step": (i32, [vp, C.c_float]),
        "storm_hip_calibrate": (i32, [vp, vp, u32, u32, u64, u32, C.POINTER(CostModel)]),
        "storm_hip_calibrate_samples": (i32, [vp, vp, u32, u32, vp, vp, vp,
                                              C.POINTER(CostModel)]),
        "storm_hip_batch_create": (i32, [vp, vp, u32, pvp]),
        "storm_hip_batch_allpairs": (i32, [vp, vp, C.POINTER(Options), C.POINTER(Stats), vp,
                                           vp]),
        "storm_hip_batch_free": (None, [vp]),
        "storm_hip_set_bb_config": (i32, [vp, C.POINTER(BBConfig)]),
        "storm_hip_get_bb_config": (i32, [vp, C.POINTER(BBConfig)]),
        "storm_hip_autotune": (i32, [vp, vp, C.POINTER(Options), C.POINTER(TuneResult)]),
        "storm_hip_corpus_from_device_lists": (i32, [vp, vp, vp, u32, u32, u32, pvp]),
        "storm_hip_stormbin_read": (i32, [C.c_char_p, u32, u32,
                                          C.POINTER(C.POINTER(C.c_uint64)),
                                          C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(u32),
                                          C.POINTER(u32)]),
        "storm_hip_corpus_from_stormbin": (i32, [vp, C.c_char_p, u32, u32, u32, pvp]),
        "storm_hip_corpus_arena_sizes": (i32, [vp, vp, C.POINTER(u64)]),
        "storm_hip_corpus_download_arena": (i32, [vp, vp, u32, vp, u64]),
        "storm_hip_allpairs_threshold": (i32, [vp, vp, C.POINTER(Options), C.c_double,
                                               C.POINTER(Stats), C.POINTER(u64)]),
        "storm_hip_result_copy_pairs": (i32, [vp, vp, u64, u64, vp]),
        "storm_hip_allpairs_xy": (i32, [vp, vp, vp, u32, C.POINTER(u64), vp]),
        "storm_hip_corpus_gen": (i32, [vp, u32, u32, u32, C.c_double, C.c_double, C.c_double,
                                       u64, u32, pvp]),
        "storm_hip_corpus_closed_form": (i32, [vp, vp, C.POINTER(u64), vp]),
        "storm_hip_calibrate_corpora": (i32, [vp, vp, u32, u32, u32, vp, vp, vp, vp,
                                              C.POINTER(CostModel)]),
        "storm_hip_allpairs_contig_noalloc": (i32, [vp, vp, u32, u32, u32, u32, vp, vp]),
        "storm_hip_selector_explore": (i32, [vp, vp, u32, u64, C.c_double, u32, vp, vp, vp,
                                             vp, C.POINTER(u32), C.POINTER(u64)]),
        "storm_hip_batch_explore": (i32, [vp, vp, u32, u64, C.c_double, u32, vp, vp, vp, vp,
                                          vp, C.POINTER(u32), vp]),
        "storm_hip_allpairs_band": (i32, [vp, vp, C.POINTER(Options), u32, u32, vp,
                                          C.POINTER(Stats)]),
        "storm_hip_allpairs_cross": (i32, [vp, vp, C.POINTER(Options), u32, vp,
                                           C.POINTER(Stats)]),
        "storm_hip_corpus_order": (i32, [vp, vp, vp]),
        "storm_hip_allpairs_threshold_prefix": (i32, [vp, vp, C.c_double, u32,
                                                      C.POINTER(Options), C.POINTER(Stats),
                                                      C.POINTER(u64), C.POINTER(PrefixInfo)]),
    }
    for name, (res, args) in sig.items():
        if not hasattr(L, name):
            continue
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _lib = L
    return L


_hiprt = None


def hip_runtime():
    """The HIP runtime libstorm_hip.so is linked against (loaded with it).
    Device buffers for the engine come from here: a second HIP runtime in the
    process (torch bundles its own) does not share devices with this one."""
    global _hiprt
    if _hiprt is None:
        lib()
        R = C.CDLL("libamdhip64.so.7")
        R.hipMalloc.argtypes, R.hipMalloc.restype = [C.POINTER(C.c_void_p), C.c_size_t], C.c_int
        R.hipFree.argtypes, R.hipFree.restype = [C.c_void_p], C.c_int
        R.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        R.hipMemcpy.restype = C.c_int
        _hiprt = R
    return _hiprt


class DeviceArray:
    """A copy of a 1-D numpy array in device memory (hipMalloc + hipMemcpy on
    the engine's HIP runtime)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.dtype, self.count = a.dtype, a.size
        self.ptr = C.c_void_p()
        R = hip_runtime()
        if R.hipMalloc(C.byref(self.ptr), max(a.nbytes, 1)) != 0:
            raise StormHipError("hipMalloc failed")
        if a.nbytes and R.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) != 0:
            self.free()
            raise StormHipError("hipMemcpy to the device failed")

    def free(self):
        if self.ptr:
            hip_runtime().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


def check(status, what=""):
    if status != 0:
        L = lib()
        raise StormHipError(f"{what}: {L.storm_hip_strerror(status).decode()} "
                            f"({status}): {L.storm_hip_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def device_count():
    n = C.c_int(0)
    check(lib().storm_hip_device_count(C.byref(n)), "device_count")
    return n.value


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().storm_hip_ctx_create(device, 0, C.byref(self.h)), "ctx_create")

    def close(self):
        if self.h:
            lib().storm_hip_ctx_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(lib().storm_hip_sync(self.h), "sync")

    # -- corpora -------------------------------------------------------------
    def upload_contig(self, words):
        words = np.ascontiguousarray(words, dtype=np.uint64)
        assert words.ndim == 2
        h = C.c_void_p()
        check(lib().storm_hip_corpus_upload_contig(self.h, _ptr(words), words.shape[0],
                                                   words.shape[1], C.byref(h)), "upload_contig")
        return Corpus(self, h, words.shape[0])

    def gen_bernoulli(self, n_rows, n_bits, p, seed):
        h = C.c_void_p()
        check(lib().storm_hip_corpus_gen_bernoulli(self.h, n_rows, n_bits, p, seed, C.byref(h)),
              "gen_bernoulli")
        return Corpus(self, h, n_rows)

    def from_lists(self, offs, pos, universe, bitmap_policy=0, on_host=False):
        """Upload CSR lists and build every representation (build_row
        semantics) on the device, or on host threads with on_host=True.
        offs: n+1 int64 offsets into pos (u32, sorted per row). bitmap_policy:
        bits 0-1 are 0 = bitmap for rows of density >= 1/64, 1 = every row,
        2 = no row; the STORM_HIP_BUILD_* bits (BUILD_*) select arenas."""
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        pos = np.ascontiguousarray(pos, dtype=np.uint32)
        n = len(offs) - 1
        flags = bitmap_policy | (BUILD_ON_HOST if on_host else 0)
        h = C.c_void_p()
        check(lib().storm_hip_corpus_from_lists(self.h, _ptr(offs), _ptr(pos), n, universe,
                                                flags, C.byref(h)), "from_lists")
        return Corpus(self, h, n)

    def from_device_lists(self, offs, pos, universe, bitmap_policy=0):
        """As from_lists for lists already in device memory: DeviceArray
        objects holding n+1 u64 offsets and the u32 positions."""
        n = offs.count - 1
        h = C.c_void_p()
        check(lib().storm_hip_corpus_from_device_lists(self.h, offs.ptr, pos.ptr, n, universe,
                                                       bitmap_policy, C.byref(h)),
              "from_device_lists")
        return Corpus(self, h, n)

    def from_stormbin(self, path, max_rows=0, stride=1, bitmap_policy=0):
        """STORMBIN file -> device corpus (records gathered and built on the
        device); corpus row k is the k-th kept (non-empty, stride-selected) row."""
        h = C.c_void_p()
        check(lib().storm_hip_corpus_from_stormbin(self.h, os.fsencode(path), max_rows, stride,
                                                   bitmap_policy, C.byref(h)), "from_stormbin")
        return Corpus(self, h, int(lib().storm_hip_corpus_rows(h)))

    def upload_arenas(self, universe, n_rows, arenas):
        """storm_hip_corpus_upload: arenas is a dict of numpy arrays keyed
        list_off/list, run_off/run_start/run_end, ewah_off/ewah,
        bitmap_off/bitmap, rank_off/rank, occ_off/occ, comp_off/comp and meta
        (8 u32 per row); a family whose offsets are missing is absent."""
        a = Arena()
        a.n_rows, a.universe = n_rows, universe
        keep = []
        types = {"list": C.c_uint32, "run_start": C.c_uint32, "run_end": C.c_uint32,
                 "comp": C.c_uint32, "meta": C.c_uint32}
        for name, _ in Arena._fields_[2:]:
            if name not in arenas or arenas[name] is None:
                continue
            ct = types.get(name, C.c_uint64)
            arr = np.ascontiguousarray(arenas[name],
                                       dtype=np.uint32 if ct is C.c_uint32 else np.uint64)
            keep.append(arr)
            setattr(a, name, arr.ctypes.data_as(C.POINTER(ct)))
        h = C.c_void_p()
        check(lib().storm_hip_corpus_upload(self.h, C.byref(a), C.byref(h)), "corpus_upload")
        return Corpus(self, h, n_rows)

    def gen_device(self, kind, n_rows, universe, p1, p2=0.0, p3=0.0, seed=1, bitmap_policy=0):
        """storm_hip_corpus_gen: a seeded workload generated on the device in
        its stored representation (Bernoulli bitmaps / lists, runs)."""
        h = C.c_void_p()
        check(lib().storm_hip_corpus_gen(self.h, GEN[kind], n_rows, universe, p1, p2, p3, seed,
                                         bitmap_policy, C.byref(h)), "corpus_gen")
        return Corpus(self, h, n_rows)

    def batch(self, corpora):
        return Batch(self, corpora)

    def allpairs_xy(self, x, y, bb="auto", matrix=True):
        """Rectangular form: (sum, counts) with counts[i, j] = |Xi n Yj| for two
        dense corpora (counts None when matrix is False)."""
        s = C.c_uint64(0)
        out = np.zeros((x.n, y.n), dtype=np.uint32) if matrix else None
        check(lib().storm_hip_allpairs_xy(self.h, x.h, y.h, BB[bb], C.byref(s), _ptr(out)),
              "allpairs_xy")
        return s.value, out

    # -- learned selector ------------------------------------------------------
    def selector_sample(self, corpus, tile=64, max_tiles=512, seed=1, repeats=2):
        F, O = mlp_shape()[0], 10
        X = np.zeros((max_tiles, F), np.float32)
        Y = np.zeros((max_tiles, O), np.float32)
        M = np.zeros((max_tiles, O), np.uint8)
        idx = np.zeros(max_tiles, np.uint32)
        n = C.c_uint32(0)
        check(lib().storm_hip_selector_sample(self.h, corpus.h, tile, max_tiles, seed, repeats,
                                              _ptr(X), _ptr(Y), _ptr(M), _ptr(idx), C.byref(n)),
              "selector_sample")
        k = n.value
        return X[:k], Y[:k], M[:k], idx[:k]

    def selector_explore(self, corpus, tile=256, seed=1, fraction=1.0):
        """One exploration call: (X, Y, M, idx, sum) with one labelled cell per
        non-empty tile pair, its cost measured in the production mix."""
        F, O = mlp_shape()[0], 10
        T = (corpus.n + tile - 1) // tile
        cap = max(1, T * (T + 1) // 2)
        X = np.zeros((cap, F), np.float32)
        Y = np.zeros((cap, O), np.float32)
        M = np.zeros((cap, O), np.uint8)
        idx = np.zeros(cap, np.uint32)
        n, s = C.c_uint32(0), C.c_uint64(0)
        check(lib().storm_hip_selector_explore(self.h, corpus.h, tile, seed, fraction, cap,
                                               _ptr(X), _ptr(Y),
                                               _ptr(M), _ptr(idx), C.byref(n), C.byref(s)),
              "selector_explore")
        k = n.value
        return X[:k], Y[:k], M[:k], idx[:k], s.value

    def selector_train(self, X, Y, M, epochs=200, batch=256, lr=3e-3, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        M = np.ascontiguousarray(M, np.uint8)
        loss = C.c_float(0)
        check(lib().storm_hip_selector_train(self.h, _ptr(X), _ptr(Y), _ptr(M), len(X), epochs,
                                             batch, lr, seed, C.byref(loss)), "selector_train")
        return loss.value

    def selector_predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros((len(X), 10), np.float32)
        check(lib().storm_hip_selector_predict(self.h, _ptr(X), len(X), _ptr(out)), "predict")
        return out

    def selector_set_params(self, params, norm=None):
        params = np.ascontiguousarray(params, np.float32)
        norm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        check(lib().storm_hip_selector_set_params(self.h, _ptr(params), _ptr(norm)), "set_params")

    def selector_get_params(self):
        F, _, _, _, P = mlp_shape()
        p = np.zeros(P, np.float32)
        nm = np.zeros(4 * F, np.float32)
        check(lib().storm_hip_selector_get_params(self.h, _ptr(p), _ptr(nm)), "get_params")
        return p, nm

    def selector_grad(self, X, Y, M):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        M = np.ascontiguousarray(M, np.uint8)
        g = np.zeros(mlp_shape()[4], np.float32)
        loss = C.c_float(0)
        check(lib().storm_hip_selector_grad(self.h, _ptr(X), _ptr(Y), _ptr(M), len(X), _ptr(g),
                                            C.byref(loss)), "selector_grad")
        return g, loss.value

    def selector_adam_step(self, lr):
        check(lib().storm_hip_selector_adam_step(self.h, lr), "adam_step")

    def calibrate(self, corpus, tile=64, max_tiles=512, seed=1, repeats=2):
        m = CostModel()
        check(lib().storm_hip_calibrate(self.h, corpus.h, tile, max_tiles, seed, repeats,
                                        C.byref(m)), "calibrate")
        return m.as_dict()

    def calibrate_samples(self, corpus, tile, Y, M, idx):
        """Fit the analytic rates from selector_sample's output (Y, M, idx)."""
        Y = np.ascontiguousarray(Y, np.float32)
        M = np.ascontiguousarray(M, np.uint8)
        idx = np.ascontiguousarray(idx, np.uint32)
        m = CostModel()
        check(lib().storm_hip_calibrate_samples(self.h, corpus.h, tile, len(idx), _ptr(Y),
                                                _ptr(M), _ptr(idx), C.byref(m)),
              "calibrate_samples")
        return m.as_dict()

    def calibrate_corpora(self, corpora, tile, Y, M, idx, cidx):
        """Fit the analytic rates from samples of several corpora (sample k
        from corpora[cidx[k]])."""
        Y = np.ascontiguousarray(Y, np.float32)
        M = np.ascontiguousarray(M, np.uint8)
        idx = np.ascontiguousarray(idx, np.uint32)
        cidx = np.ascontiguousarray(cidx, np.uint32)
        hs = (C.c_void_p * len(corpora))(*[c.h for c in corpora])
        m = CostModel()
        check(lib().storm_hip_calibrate_corpora(self.h, hs, len(corpora), tile, len(idx), _ptr(Y),
                                                _ptr(M), _ptr(idx), _ptr(cidx), C.byref(m)),
              "calibrate_corpora")
        return m.as_dict()

    def set_cost_model(self, rates, ns_fixed=2.0, ns_per_occ_word=0.30, hot=0.0, cold=0.0,
                       bb_zonemap=1.0):
        m = CostModel()
        for i, r in enumerate(rates):
            m.ns_per_unit[i] = r
        m.ns_fixed, m.ns_per_occ_word, m.ns_probe_hot, m.ns_probe_cold = ns_fixed, ns_per_occ_word, hot, cold
        m.bb_zonemap = bb_zonemap
        check(lib().storm_hip_set_cost_model(self.h, C.byref(m)), "set_cost_model")

    def get_cost_model(self):
        m = CostModel()
        check(lib().storm_hip_get_cost_model(self.h, C.byref(m)), "get_cost_model")
        return m

    def put_cost_model(self, m):
        """Install a CostModel returned by get_cost_model."""
        check(lib().storm_hip_set_cost_model(self.h, C.byref(m)), "set_cost_model")

    # -- launch shapes -----------------------------------------------------------
    def set_bb_config(self, stage_words=8, band=4, blocks_per_cu=2, waves=8):
        c = BBConfig(stage_words, band, blocks_per_cu, waves)
        check(lib().storm_hip_set_bb_config(self.h, C.byref(c)), "set_bb_config")

    def get_bb_config(self):
        c = BBConfig()
        check(lib().storm_hip_get_bb_config(self.h, C.byref(c)), "get_bb_config")
        return c.as_dict()

    def autotune(self, corpus, policy="analytic", bb="auto", tile=64, use_graph=False):
        o = Options()
        o.policy, o.out_mode, o.bb_kernel = POLICY[policy], OUT["sum"], BB[bb]
        o.tile, o.use_graph = tile, int(use_graph)
        r = TuneResult()
        check(lib().storm_hip_autotune(self.h, corpus.h, C.byref(o), C.byref(r)), "autotune")
        return r.as_dict()


class Corpus:
    def __init__(self, ctx, h, n):
        self.ctx, self.h, self.n = ctx, h, n

    def free(self):
        if self.h:
            lib().storm_hip_corpus_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    @property
    def n_pairs(self):
        return self.n * (self.n - 1) // 2

    def device_bytes(self):
        return int(lib().storm_hip_corpus_device_bytes(self.h))

    def download_words(self, row0, n_rows, n_words):
        out = np.zeros((n_rows, n_words), dtype=np.uint64)
        check(lib().storm_hip_corpus_download_words(self.ctx.h, self.h, row0, n_rows, _ptr(out)),
              "download_words")
        return out

    def materialize_bitmaps(self):
        check(lib().storm_hip_corpus_materialize_bitmaps(self.ctx.h, self.h), "materialize")

    def allpairs(self, policy="all_bitmap", out="sum", bb="auto", tile=64, fixed_cell=0,
                 use_graph=False):
        o = Options()
        o.policy, o.out_mode, o.bb_kernel = POLICY[policy], OUT[out], BB[bb]
        o.tile, o.fixed_cell, o.use_graph = tile, fixed_cell, int(use_graph)
        st = Stats()
        check(lib().storm_hip_allpairs(self.ctx.h, self.h, C.byref(o), C.byref(st)), "allpairs")
        return st.as_dict()

    def allpairs_threshold(self, t, policy="analytic", bb="auto", tile=64, fixed_cell=0,
                           use_graph=False):
        """Thresholded mode: (stats, pairs) with pairs an (n, 3) uint32 array of
        (i, j, |Xi n Xj|), i < j, sorted, for every pair whose Jaccard >= t."""
        o = Options()
        o.policy, o.out_mode, o.bb_kernel = POLICY[policy], OUT["sum"], BB[bb]
        o.tile, o.fixed_cell, o.use_graph = tile, fixed_cell, int(use_graph)
        st = Stats()
        n = C.c_uint64(0)
        check(lib().storm_hip_allpairs_threshold(self.ctx.h, self.h, C.byref(o), t, C.byref(st),
                                                 C.byref(n)), "allpairs_threshold")
        out = np.zeros((n.value, 3), dtype=np.uint32)
        check(lib().storm_hip_result_copy_pairs(self.ctx.h, self.h, 0, n.value, _ptr(out)),
              "copy_pairs")
        return st.as_dict(), out

    def allpairs_threshold_prefix(self, t, mode="gated", policy="analytic", tile=256):
        """Thresholded mode by prefix-filtered candidates
        (storm_hip_allpairs_threshold_prefix): (stats, pairs, info) with pairs
        as allpairs_threshold's and info the gate's estimate and candidate
        counts; mode "gated", "prefix" or "exact"."""
        o = Options()
        o.policy, o.out_mode, o.bb_kernel = POLICY[policy], OUT["sum"], BB["auto"]
        o.tile = tile
        st = Stats()
        info = PrefixInfo()
        n = C.c_uint64(0)
        check(lib().storm_hip_allpairs_threshold_prefix(
            self.ctx.h, self.h, t, {"gated": 0, "prefix": 1, "exact": 2}[mode], C.byref(o),
            C.byref(st), C.byref(n), C.byref(info)), "allpairs_threshold_prefix")
        out = np.zeros((n.value, 3), dtype=np.uint32)
        check(lib().storm_hip_result_copy_pairs(self.ctx.h, self.h, 0, n.value, _ptr(out)),
              "copy_pairs")
        return st.as_dict(), out, {k: getattr(info, k) for k, _ in info._fields_}

    def allpairs_band(self, band, policy="analytic", order="density", bb="auto", tile=64,
                      fixed_cell=0, use_graph=False):
        """(stats, counts): tile row `band` of the upper triangle in the density
        or identity order, counts[(i - band*tile), j] for positions i < j."""
        o = Options()
        o.policy, o.out_mode, o.bb_kernel = POLICY[policy], OUT["sum"], BB[bb]
        o.tile, o.fixed_cell, o.use_graph = tile, fixed_cell, int(use_graph)
        rows = max(0, min(tile, self.n - band * tile))
        out = np.zeros((rows, self.n), dtype=np.uint32)
        st = Stats()
        check(lib().storm_hip_allpairs_band(self.ctx.h, self.h, C.byref(o),
                                            {"density": 0, "identity": 1}[order], band,
                                            _ptr(out), C.byref(st)), "allpairs_band")
        return st.as_dict(), out

    def allpairs_cross(self, split, policy="analytic", bb="auto", tile=64, fixed_cell=0,
                       matrix=True, use_graph=False):
        """(stats, counts): rows [0, split) against rows [split, n); counts is
        split x (n - split), or None with matrix=False (sum only)."""
        o = Options()
        o.policy, o.out_mode, o.bb_kernel = POLICY[policy], OUT["sum"], BB[bb]
        o.tile, o.fixed_cell, o.use_graph = tile, fixed_cell, int(use_graph)
        out = np.zeros((split, self.n - split), dtype=np.uint32) if matrix else None
        st = Stats()
        check(lib().storm_hip_allpairs_cross(self.ctx.h, self.h, C.byref(o), split, _ptr(out),
                                             C.byref(st)), "allpairs_cross")
        return st.as_dict(), out

    def order(self):
        """Density order: perm[k] = the row at position k."""
        out = np.empty(self.n, dtype=np.uint32)
        check(lib().storm_hip_corpus_order(self.ctx.h, self.h, _ptr(out)), "corpus_order")
        return out

    def tri(self, offset=0, count=None):
        if count is None:
            count = self.n_pairs - offset
        out = np.empty(count, dtype=np.uint32)
        check(lib().storm_hip_result_copy_tri(self.ctx.h, self.h, offset, count, _ptr(out)),
              "copy_tri")
        return out

    def tri_to_original(self, t):
        """A density-packed triangle (out="tri_density") rearranged into the
        original row order of out="tri" (host-side, for checking)."""
        n = self.n
        perm = self.order().astype(np.int64)
        iu, ju = np.triu_indices(n, 1)
        a, b = perm[iu], perm[ju]
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        dst = lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)
        out = np.empty_like(t)
        out[dst] = t
        return out

    def row_sums(self):
        out = np.empty(self.n, dtype=np.uint64)
        check(lib().storm_hip_result_row_sums(self.ctx.h, self.h, _ptr(out)), "row_sums")
        return out

    def closed_form(self, row_sums=True):
        """(total, row_sums): sum_b C(c_b, 2) and sum_{b in Xi} (c_b - 1) from
        the column counts, computed on the device without any cell."""
        t = C.c_uint64(0)
        rs = np.zeros(self.n, dtype=np.uint64) if row_sums else None
        check(lib().storm_hip_corpus_closed_form(self.ctx.h, self.h, C.byref(t), _ptr(rs)),
              "closed_form")
        return t.value, rs

    def last_decisions(self, n_tiles):
        out = np.empty(n_tiles, dtype=np.uint8)
        check(lib().storm_hip_last_decisions(self.ctx.h, self.h, _ptr(out), n_tiles),
              "last_decisions")
        return out

    def arenas(self):
        """Every arena of a general corpus as numpy arrays keyed by name."""
        sizes = (C.c_uint64 * len(ARENAS))()
        check(lib().storm_hip_corpus_arena_sizes(self.ctx.h, self.h, sizes), "arena_sizes")
        out = {}
        for k, (name, dt) in enumerate(ARENAS):
            a = np.empty(int(sizes[k]) // np.dtype(dt).itemsize, dtype=dt)
            check(lib().storm_hip_corpus_download_arena(self.ctx.h, self.h, k, _ptr(a),
                                                        int(sizes[k])), f"download {name}")
            out[name] = a
        return out


class Batch:
    """Independent general corpora run through one set of launches
    (storm_hip_batch_*). The corpora must outlive the batch."""

    def __init__(self, ctx, corpora):
        self.ctx, self.corpora = ctx, list(corpora)
        hs = (C.c_void_p * max(1, len(self.corpora)))(*[c.h for c in self.corpora])
        self.h = C.c_void_p()
        check(lib().storm_hip_batch_create(ctx.h, hs, len(self.corpora), C.byref(self.h)),
              "batch_create")

    def free(self):
        if self.h:
            lib().storm_hip_batch_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def explore(self, tile=256, seed=1, fraction=1.0):
        """Exploration over the batch: (X, Y, M, idx, corpus index, sums)."""
        F, O = mlp_shape()[0], 10
        cap = 0
        for c in self.corpora:
            T = (c.n + tile - 1) // tile
            cap += T * (T + 1) // 2
        cap = max(cap, 1)
        X = np.zeros((cap, F), np.float32)
        Y = np.zeros((cap, O), np.float32)
        M = np.zeros((cap, O), np.uint8)
        idx = np.zeros(cap, np.uint32)
        cidx = np.zeros(cap, np.uint32)
        sums = np.zeros(max(1, len(self.corpora)), np.uint64)
        n = C.c_uint32(0)
        check(lib().storm_hip_batch_explore(self.ctx.h, self.h, tile, seed, fraction, cap,
                                            _ptr(X), _ptr(Y),
                                            _ptr(M), _ptr(idx), _ptr(cidx), C.byref(n),
                                            _ptr(sums)), "batch_explore")
        k = n.value
        return X[:k], Y[:k], M[:k], idx[:k], cidx[:k], sums[:len(self.corpora)]

    def allpairs(self, policy="analytic", out="sum", bb="auto", tile=64, fixed_cell=0,
                 use_graph=False):
        """Returns (stats dict, per-corpus sums, per-corpus latency in ms)."""
        o = Options()
        o.policy, o.out_mode, o.bb_kernel = POLICY[policy], OUT[out], BB[bb]
        o.tile, o.fixed_cell, o.use_graph = tile, fixed_cell, int(use_graph)
        st = Stats()
        k = len(self.corpora)
        sums = np.zeros(max(1, k), dtype=np.uint64)
        lat = np.zeros(max(1, k), dtype=np.float64)
        check(lib().storm_hip_batch_allpairs(self.ctx.h, self.h, C.byref(o), C.byref(st),
                                             _ptr(sums), _ptr(lat)), "batch_allpairs")
        return st.as_dict(), sums[:k], lat[:k]


def tri_index(i, j, n):
    return i * n - i * (i + 1) // 2 + (j - i - 1)


def mlp_shape():
    v = [C.c_int(0) for _ in range(5)]
    check(lib().storm_hip_mlp_shape(*[C.byref(x) for x in v]), "mlp_shape")
    return tuple(x.value for x in v)


GEN = {"zipf": 0, "runs": 1, "bernoulli": 2}


def gen_corpus(kind, n_rows, universe, p1, p2=0.0, p3=0.0, seed=1):
    """Seeded synthetic rows (storm_hip_gen_corpus) as numpy CSR copies."""
    offs_p = C.POINTER(C.c_uint64)()
    pos_p = C.POINTER(C.c_uint32)()
    total = C.c_uint64(0)
    check(lib().storm_hip_gen_corpus(GEN[kind], n_rows, universe, p1, p2, p3, seed,
                                     C.byref(offs_p), C.byref(pos_p), C.byref(total)),
          "gen_corpus")
    try:
        offs = np.ctypeslib.as_array(offs_p, shape=(n_rows + 1,)).copy()
        pos = (np.ctypeslib.as_array(pos_p, shape=(total.value,)).copy() if total.value
               else np.zeros(0, np.uint32))
    finally:
        lib().storm_hip_gen_free(C.cast(offs_p, C.c_void_p))
        lib().storm_hip_gen_free(C.cast(pos_p, C.c_void_p))
    return offs, pos


def stormbin_read(path, max_rows=0, stride=1):
    """STORMBIN file -> (offs, pos, universe) host CSR (storm_hip_stormbin_read:
    the reference loader's row selection)."""
    offs_p = C.POINTER(C.c_uint64)()
    pos_p = C.POINTER(C.c_uint32)()
    n, u = C.c_uint32(0), C.c_uint32(0)
    check(lib().storm_hip_stormbin_read(os.fsencode(path), max_rows, stride, C.byref(offs_p),
                                        C.byref(pos_p), C.byref(n), C.byref(u)),
          "stormbin_read")
    try:
        offs = np.ctypeslib.as_array(offs_p, shape=(n.value + 1,)).copy()
        total = int(offs[-1])
        pos = (np.ctypeslib.as_array(pos_p, shape=(total,)).copy() if total
               else np.zeros(0, np.uint32))
    finally:
        lib().storm_hip_gen_free(C.cast(offs_p, C.c_void_p))
        lib().storm_hip_gen_free(C.cast(pos_p, C.c_void_p))
    return offs, pos, u.value


def write_stormbin(path, offs, pos, universe):
    """Writes CSR rows as STORMBIN (tools/sets2bin.py:23-25 layout: magic,
    <IIII> version 1, n_rows, n_bits, 0, then per row <I> count + u32 positions)."""
    offs = np.asarray(offs, dtype=np.uint64)
    pos = np.asarray(pos, dtype=np.uint32)
    n = len(offs) - 1
    counts = np.diff(offs).astype(np.uint32)
    # Interleave counts and positions in one u32 buffer.
    buf = np.empty(n + len(pos), dtype="<u4")
    heads = (np.arange(n, dtype=np.uint64) + offs[:-1]).astype(np.int64)
    buf[heads] = counts
    mask = np.ones(len(buf), dtype=bool)
    mask[heads] = False
    buf[mask] = pos
    with open(path, "wb") as f:
        f.write(b"STORMBIN")
        f.write(np.array([1, n, universe, 0], dtype="<u4").tobytes())
        f.write(buf.tobytes())


# ---- the batched-instances workload ----------------------------------------
_GOLDEN = 0x9E3779B97F4A7C15
_MASK64 = (1 << 64) - 1


class _Stream:
    """splitmix64 stream (kernels/storm_gen.h:84-102): next() and unit()."""

    def __init__(self, seed):
        self.s = (seed if seed else _GOLDEN) & _MASK64

    def next(self):
        self.s = (self.s + _GOLDEN) & _MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def unit(self):
        return (self.next() >> 11) * 2.0 ** -53


# Representation planning for the batched workload: a bitmap of every row is
# added to an instance held as lists or runs when the matrix-core B x B is
# predicted faster than the instance's list or run cell. Rates measured on
# MI355X: B x B at about 3e15 (int8) / 6e15 (fp4) bit-pair operations per
# second, the S x S hash join at 6e-13 s and the R x R interval join at
# 1e-12 s per pair and list element (run).
BB_RATE_I8, BB_RATE_FP4 = 3.0e15, 6.0e15
SS_S_PER_PAIR_ITEM, RR_S_PER_PAIR_ITEM = 6e-13, 1e-12


def batched_instances(count=512, seed=5, plan_bitmaps=True, bitmap_cap_bytes=64 << 30):
    """The independent all-pairs problems of the batched workload, drawn in
    order from Rng(seed): instance k has N uniform in [64, 4096], universe
    log-uniform in [2e5, 1.3e8], density log-uniform in [1e-6, 0.17];
    even instances are i.i.d. Bernoulli rows, odd ones clustered-run rows
    (Poisson runs of geometric length L = min(2000, d*U/20), at least 1, so
    that lambda * L = d * U, plus 5% extra singletons). Each entry is a dict
    with the generator call that builds it on the device (gen_device), its
    stored size estimate in bytes and whether every row also gets a bitmap
    (plan_bitmaps: when the B x B is predicted faster than the list or run
    cell and the bitmaps take at most bitmap_cap_bytes)."""
    import math
    g = _Stream(seed)
    out = []
    for k in range(count):
        n = 64 + int(g.unit() * 4033)
        u = int(math.exp(math.log(2e5) + g.unit() * (math.log(1.3e8) - math.log(2e5))))
        d = math.exp(math.log(1e-6) + g.unit() * (math.log(0.17) - math.log(1e-6)))
        sk = g.next() & 0xFFFFFFFF
        nw = (u + 63) // 64
        pairs = n * (n - 1) // 2
        bm_bytes = n * ((nw + 63) // 64 * 64) * 8
        t_bb = pairs * 2.0 * u / (BB_RATE_FP4 if nw < (1 << 18) else BB_RATE_I8)
        if k % 2 == 0:
            kind = "bernoulli"
            if d * 32 >= 1:      # bitmaps only
                est, flags, t_other = bm_bytes, 0, t_bb
            else:                # lists; bitmaps where a row's density is >= 1/64
                est = 4 * n * u * d + (bm_bytes if d * 64 >= 1 else 0)
                flags, t_other = 0, pairs * 2.0 * d * u * SS_S_PER_PAIR_ITEM
            args = ["bernoulli", n, u, d, 0.0, 0.0, sk, flags]
        else:
            kind = "runs"
            L = max(1.0, min(2000.0, d * u / 20.0))
            lam = max(d * u / L, 1e-3)
            runs = lam * (1 + 0.05 * L)
            est = 8 * n * (runs + 1)
            t_other = pairs * 2.0 * runs * RR_S_PER_PAIR_ITEM
            args = ["runs", n, u, lam, L, 0.05, sk, 2]   # runs only
        bitmaps = False
        if plan_bitmaps and args[7] != 1 and t_bb < t_other and bm_bytes <= bitmap_cap_bytes \
                and not (kind == "bernoulli" and d * 32 >= 1):
            args[7] = 1                                  # a bitmap for every row
            est += bm_bytes if not (kind == "bernoulli" and d * 64 >= 1) else 0
            bitmaps = True
        out.append({"k": k, "n": n, "universe": u, "density": d, "kind": kind, "seed": sk,
                    "gen": tuple(args), "est_bytes": int(est) + 64 * n + (1 << 20),
                    "pairs": pairs, "bitmaps_added": bitmaps,
                    "predicted_s": min(t_bb, t_other) if bitmaps else t_other})
    return out


def gen_instance(ctx, inst):
    kind, n, u, p1, p2, p3, seed, flags = inst["gen"]
    return ctx.gen_device(kind, n, u, p1, p2, p3, seed, bitmap_policy=flags)
