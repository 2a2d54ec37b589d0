"""ctypes loader for the CPU checkers. TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this.
  liboracle.so          plain-C restatement of the reference (storm_oracle.c)
  _ref/libstormref_*.so the reference itself, compiled from its sources
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_orc = None
_ref = None


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def oracle():
    global _orc
    if _orc is not None:
        return _orc
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
    L = C.CDLL(ORACLE_SO)
    vp, u32, u64, i32, dbl = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    sig = {
        "orc_rng_next": (u64, [vp]),
        "orc_rng_below": (u32, [vp, u32]),
        "orc_rng_seed": (u64, [u64]),
        "orc_gen_bernoulli_rows": (None, [u64, dbl, u32, u32, u32, vp]),
        "orc_popcount_bitloop": (u32, [u64]),
        "orc_intersect_lists": (u64, [vp, u32, u32, vp, u32, u32]),
        "orc_intersect_merge": (u64, [vp, u32, vp, u32]),
        "orc_intersect_words": (u64, [vp, vp, u32]),
        "orc_build_rank": (None, [vp, u32, vp]),
        "orc_build_occ": (None, [vp, u32, u32, vp]),
        "orc_build_ewah": (u32, [vp, u32, vp]),
        "orc_ewah_decode": (u32, [vp, u32, vp]),
        "orc_allpairs_lists": (u64, [vp, vp, u32, vp, i32]),
        "orc_allpairs_words": (u64, [vp, u32, u32, vp, i32]),
        "orc_column_counts_words": (None, [vp, u32, u32, vp]),
        "orc_total_from_columns": (u64, [vp, u64]),
        "orc_column_closed_forms_words": (None, [vp, u32, u32, vp, vp, i32]),
        "orc_band_words": (u64, [vp, u32, u32, u32, u32, vp, i32]),
        "orc_row_sums_words": (None, [vp, u32, u32, vp, vp]),
        "orc_row_sums_from_tri": (None, [vp, u32, vp]),
        "orc_select_pairing": (i32, [vp, vp, vp, vp]),
        "orc_predict": (dbl, [vp, i32, vp, vp, dbl, dbl, dbl]),
        "orc_default_model": (None, [vp]),
        "orc_tile_meta": (None, [vp, vp, u32, u32, vp]),
        "orc_density_order": (None, [vp, u32, vp]),
        "orc_mlp_forward_backward": (dbl, [i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp,
                                           vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "orc_adam_step": (None, [i32, vp, vp, vp, vp, i32, dbl, dbl, dbl, dbl]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    _orc = L
    return L


def _cpu_has_avx512():
    """CPUID through liboracle (built without ISA flags): AVX-512 F, BW, VL
    and VPOPCNTDQ all present."""
    L = oracle()
    L.orc_cpu_has_avx512_vpopcntdq.restype = C.c_int
    return bool(L.orc_cpu_has_avx512_vpopcntdq())


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def ref_isa():
    return "avx512 (F/BW/VL/VPOPCNTDQ)" if _cpu_has_avx512() else "x86-64-v3 (AVX2)"


def ref_path():
    name = "libstormref_avx512.so" if _cpu_has_avx512() else "libstormref_x86v3.so"
    return os.path.join(REF_DIR, name)


def ref_available():
    return os.path.exists(ref_path())


def reference():
    """The compiled StormBitmaps reference (oracle/_ref), with ref_shim.cpp."""
    global _ref
    if _ref is not None:
        return _ref
    path = ref_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle ref` where the reference "
                           "sources are present")
    L = C.CDLL(path, mode=C.RTLD_LOCAL)
    vp, u32, u64, i32, dbl = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    sig = {
        "ref_rng_next": (None, [u64, u32, vp]),
        "ref_rng_below": (None, [u64, u32, u32, vp]),
        "ref_build_row": (i32, [vp, u32, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "ref_build_rank": (None, [vp, u32, vp]),
        "ref_select_pairing": (i32, [vp, vp, vp]),
        "ref_predict": (dbl, [i32, vp, vp]),
        "ref_allpairs_sum": (u64, [vp, vp, u32, u32, i32, u32, vp, vp, vp]),
        "ref_allpairs_tiles_tri": (u64, [vp, vp, u32, u32, u32, vp]),
        "ref_generate": (u64, [u32, u32, dbl, i32, i32, u64, vp, vp]),
        "ref_bb_dense_allpairs": (u64, [vp, u32, u32, i32, vp]),
        "ref_time_allpairs": (u64, [vp, vp, u32, u32, i32, u32, i32, vp]),
        "ref_time_contig": (u64, [vp, u32, u32, u32, vp]),
        "STORM_contig_new": (vp, [C.c_size_t]),
        "STORM_contig_free": (None, [vp]),
        "STORM_contig_add": (i32, [vp, vp, u32]),
        "STORM_contig_clear": (i32, [vp]),
        "STORM_contig_pairw_intersect_cardinality": (u64, [vp]),
        "STORM_contig_pairw_intersect_cardinality_blocked": (u64, [vp, u32]),
        "STORM_new": (vp, []),
        "STORM_free": (None, [vp]),
        "STORM_add": (i32, [vp, vp, u32]),
        "STORM_clear": (i32, [vp]),
        "STORM_pairw_intersect_cardinality": (u64, [vp]),
        "STORM_pairw_intersect_cardinality_blocked": (u64, [vp, u32]),
        "STORM_intersect_cardinality_square": (u64, [vp, vp]),
        "STORM_serialized_size": (u64, [vp]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    _ref = L
    return L


# ---- convenience wrappers -------------------------------------------------

def bernoulli_rows(seed, p, n_bits, row0, n_rows):
    nw = (n_bits + 63) // 64
    out = np.zeros((n_rows, nw), dtype=np.uint64)
    oracle().orc_gen_bernoulli_rows(seed, p, n_bits, row0, n_rows, _p(out))
    return out


def allpairs_words(words, tri=True, threads=8):
    words = np.ascontiguousarray(words, dtype=np.uint64)
    n, nw = words.shape
    out = np.zeros(n * (n - 1) // 2, dtype=np.uint32) if tri else None
    s = oracle().orc_allpairs_words(_p(words), n, nw, _p(out), threads)
    return s, out


def allpairs_lists(offs, pos, tri=True, threads=8):
    offs32 = np.ascontiguousarray(offs, dtype=np.uint32)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    n = len(offs32) - 1
    out = np.zeros(n * (n - 1) // 2, dtype=np.uint32) if tri else None
    s = oracle().orc_allpairs_lists(_p(offs32), _p(pos), n, _p(out), threads)
    return s, out


def column_closed_forms(words, threads=1):
    """(total, row_sums) from column counts -- exact, O(N*m)."""
    words = np.ascontiguousarray(words, dtype=np.uint64)
    n, nw = words.shape
    col = np.zeros(nw * 64, dtype=np.uint32)
    L = oracle()
    rs = np.zeros(n, dtype=np.uint64)
    if threads > 1:
        L.orc_column_closed_forms_words(_p(words), n, nw, _p(col), _p(rs), threads)
        return L.orc_total_from_columns(_p(col), nw * 64), rs
    L.orc_column_counts_words(_p(words), n, nw, _p(col))
    total = L.orc_total_from_columns(_p(col), nw * 64)
    L.orc_row_sums_words(_p(words), n, nw, _p(col), _p(rs))
    return total, rs


def band_words(words, r0, r1, threads=16):
    """Triangle entries of rows [r0, r1) (every j > i), in triangle order."""
    words = np.ascontiguousarray(words, dtype=np.uint64)
    n, nw = words.shape
    r1 = min(r1, n)
    cnt = sum(n - i - 1 for i in range(r0, r1))
    out = np.zeros(max(cnt, 1), dtype=np.uint32)
    s = oracle().orc_band_words(_p(words), n, nw, r0, r1, _p(out), threads)
    return s, out[:cnt]


def lists_closed_forms(offs, pos, universe):
    """(total, row_sums) from column counts over CSR lists."""
    offs = np.asarray(offs, dtype=np.int64)
    pos = np.asarray(pos, dtype=np.int64)
    col = np.bincount(pos, minlength=universe).astype(np.int64)
    total = int((col * (col - 1) // 2).sum())
    contrib = col[pos] - 1
    cs = np.concatenate([[0], np.cumsum(contrib)])
    rs = (cs[offs[1:]] - cs[offs[:-1]]).astype(np.uint64)
    return total, rs


class OrcRow(C.Structure):
    """orc_row (storm_oracle.h)."""
    _fields_ = [("bitmap", C.POINTER(C.c_uint64)), ("n_words", C.c_uint32),
                ("rank", C.POINTER(C.c_uint64)), ("n_rank", C.c_uint32),
                ("occ", C.POINTER(C.c_uint64)), ("n_occ", C.c_uint32),
                ("list", C.POINTER(C.c_uint32)), ("n_list", C.c_uint32),
                ("run_start", C.POINTER(C.c_uint32)), ("run_end", C.POINTER(C.c_uint32)),
                ("n_runs", C.c_uint32),
                ("ewah", C.POINTER(C.c_uint64)), ("n_ewah", C.c_uint32),
                ("comp_list", C.POINTER(C.c_uint32)), ("n_comp", C.c_uint32),
                ("comp_start", C.POINTER(C.c_uint32)), ("comp_end", C.POINTER(C.c_uint32)),
                ("n_comp_runs", C.c_uint32),
                ("meta", C.c_uint32 * 6), ("has_rank", C.c_uint8), ("has_occ", C.c_uint8)]


def build_row(pos, u):
    """orc_build_row (build_row, storm_repr.cpp:102-207) as numpy arrays."""
    L = oracle()
    L.orc_build_row.argtypes = [C.POINTER(OrcRow), C.c_void_p, C.c_uint32, C.c_uint32]
    L.orc_build_row.restype = C.c_int
    L.orc_free_row.argtypes = [C.POINTER(OrcRow)]
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    r = OrcRow()
    if L.orc_build_row(C.byref(r), _p(pos), len(pos), u) != 0:
        raise ValueError("orc_build_row failed")

    def arr(p, n, dt):
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy() if n else np.zeros(0, dt)

    out = dict(meta=np.array(list(r.meta), np.uint32), bitmap=arr(r.bitmap, r.n_words, np.uint64),
               rank=arr(r.rank, r.n_rank, np.uint64), occ=arr(r.occ, r.n_occ, np.uint64),
               run_start=arr(r.run_start, r.n_runs, np.uint32),
               run_end=arr(r.run_end, r.n_runs, np.uint32),
               ewah=arr(r.ewah, r.n_ewah, np.uint64), comp=arr(r.comp_list, r.n_comp, np.uint32),
               has_rank=r.has_rank, has_occ=r.has_occ)
    L.orc_free_row(C.byref(r))
    return out


def ref_generate(n_rows, universe, density, structure, spectrum, seed):
    """Corpus from the reference generator (kernels/storm_gen.cpp:173-232)."""
    L = reference()
    tot = L.ref_generate(n_rows, universe, density, structure, spectrum, seed, None, None)
    offs = np.zeros(n_rows + 1, dtype=np.uint32)
    pos = np.zeros(max(tot, 1), dtype=np.uint32)
    L.ref_generate(n_rows, universe, density, structure, spectrum, seed, _p(offs), _p(pos))
    return offs.astype(np.uint64), pos[:tot]


def ref_build_row(pos, u):
    """The reference's build_row (kernels/storm_repr.cpp:102-207) arenas of one
    row: bitmap, rank, occ, runs, ewah, complement and the six meta integers,
    plus has_rank / has_occ in sizes[7:9]."""
    L = reference()
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    sizes = np.zeros(9, np.uint32)
    meta = np.zeros(6, np.uint32)
    L.ref_build_row(_p(pos), len(pos), u, _p(sizes), _p(meta), None, None, None, None, None,
                    None, None)
    nw, nrank, nocc, nruns, newah, ncomp = (int(x) for x in sizes[:6])
    bm = np.zeros(max(nw, 1), np.uint64)
    rank = np.zeros(max(nrank, 1), np.uint64)
    occ = np.zeros(max(nocc, 1), np.uint64)
    rs = np.zeros(max(nruns, 1), np.uint32)
    re_ = np.zeros(max(nruns, 1), np.uint32)
    ew = np.zeros(max(newah, 1), np.uint64)
    comp = np.zeros(max(ncomp, 1), np.uint32)
    L.ref_build_row(_p(pos), len(pos), u, _p(sizes), _p(meta), _p(bm), _p(rank), _p(occ),
                    _p(rs), _p(re_), _p(ew), _p(comp))
    return dict(sizes=sizes, meta=meta, bitmap=bm[:nw], rank=rank[:nrank], occ=occ[:nocc],
                run_start=rs[:nruns], run_end=re_[:nruns], ewah=ew[:newah], comp=comp[:ncomp])
