"""Representation building on the device (build_dev.hip): every arena against
the host builder and, row by row, against the oracle's build_row
(storm_repr.cpp:102-207); invalid rows rejected; STORMBIN ingest and
device-resident lists giving the same corpus as host lists."""
import numpy as np
import pytest

import corpora
import oracle_py as O

pytestmark = pytest.mark.gpu

BUILD_ALL = 4 | 8 | 16
NO_BITMAP = 0xFFFFFFFF


def _generated(sh):
    """Larger rows than the edge sets: skewed, run-heavy, dense (complements,
    all-one words), each with empty rows mixed in."""
    out = []
    offs, pos = sh.gen_corpus("zipf", 600, 200_003, 1.1, 20_000.0, 0.0, seed=31)
    out.append(("zipf", offs, pos, 200_003))
    offs, pos = sh.gen_corpus("runs", 300, 1_000_000, 6.0, 900.0, 0.05, seed=32)
    out.append(("runs", offs, pos, 1_000_000))
    offs, pos = sh.gen_corpus("bernoulli", 200, 9_000, 0.7, seed=33)
    out.append(("bernoulli0.7", offs, pos, 9_000))
    rows = corpora.edge_rows(70_001) + [np.zeros(0, np.uint32)] * 3
    rows += [np.arange(640, 70_001, dtype=np.uint32), np.arange(0, 3_200, dtype=np.uint32)]
    offs, pos = corpora.csr(rows)
    out.append(("edge+full", offs, pos, 70_001))
    return out


def _cases(sh):
    out = [(name,) + corpora.csr(rows) + (u,) for name, rows, u in corpora.all_corpora()]
    return out + _generated(sh)


def _same_arenas(a, b, what):
    assert a.keys() == b.keys()
    for k in a:
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), (what, k, a[k].size,
                                                                         b[k].size)


def test_device_build_matches_host_builder(ctx, sh):
    for name, offs, pos, u in _cases(sh):
        print("case", name, flush=True)
        for flags in (BUILD_ALL, 1 | BUILD_ALL, 2):
            with ctx.from_lists(offs, pos, u, bitmap_policy=flags) as d, \
                    ctx.from_lists(offs, pos, u, bitmap_policy=flags, on_host=True) as h:
                _same_arenas(d.arenas(), h.arenas(), (name, flags))


def test_representation_subsets(ctx, sh):
    name, offs, pos, u = _generated(sh)[1]
    for flags in (4, 8, 16, 4 | 16):
        print("flags", flags, flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=flags) as d, \
                ctx.from_lists(offs, pos, u, bitmap_policy=flags, on_host=True) as h:
            A = d.arenas()
            _same_arenas(A, h.arenas(), flags)
            if not flags & 4:
                assert A["run_start"].size == 0 and not A["run_off"].any()
            if not flags & 8:
                assert A["ewah"].size == 0 and A["mk_word"].size == 0


def test_device_build_matches_oracle_rows(ctx, sh):
    cases = [c for c in _cases(sh) if c[0].startswith("edge") or "-d0.3-" in c[0]
             or c[0] in ("zipf", "bernoulli0.7")]
    for name, offs, pos, u in cases:
        print("case", name, flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as d:
            A = d.arenas()
        nw = (u + 63) // 64
        ld = A["bitmap"].size // max(1, int((A["bm_row"] != NO_BITMAP).sum()))
        for r in range(len(offs) - 1):
            row = pos[offs[r]:offs[r + 1]]
            want = O.build_row(row, u)
            m = A["meta"][8 * r:8 * r + 8]
            assert np.array_equal(m[:6], want["meta"]), (name, r)
            assert m[6] == want["has_rank"] and m[7] == want["has_occ"], (name, r)
            rs = slice(int(A["run_off"][r]), int(A["run_off"][r + 1]))
            assert np.array_equal(A["run_start"][rs], want["run_start"]), (name, r)
            assert np.array_equal(A["run_end"][rs], want["run_end"]), (name, r)
            es = slice(int(A["ewah_off"][r]), int(A["ewah_off"][r + 1]))
            assert np.array_equal(A["ewah"][es], want["ewah"]), (name, r)
            cs = slice(int(A["comp_off"][r]), int(A["comp_off"][r + 1]))
            assert np.array_equal(A["comp"][cs], want["comp"]), (name, r)
            # marker table: decoded start word and buffer offset of each marker
            ew = want["ewah"]
            q, w, mk_w, mk_p = 0, 0, [], []
            while q < len(ew):
                mk_w.append(w)
                mk_p.append(q)
                fl, ll = (int(ew[q]) >> 1) & 0x7FFFFFFF, int(ew[q]) >> 32
                w += fl + ll
                q += 1 + ll
            ms = slice(int(A["mk_off"][r]), int(A["mk_off"][r + 1]))
            assert A["mk_word"][ms].tolist() == mk_w and A["mk_pos"][ms].tolist() == mk_p
            b = int(A["bm_row"][r])
            if b != NO_BITMAP:
                assert np.array_equal(A["bitmap"][b * ld:b * ld + nw], want["bitmap"]), (name, r)
                assert not A["bitmap"][b * ld + nw:(b + 1) * ld].any()


@pytest.mark.parametrize("row", [[5, 3], [3, 3], [0, 1, 100], [100]])
def test_device_build_rejects_invalid_rows(ctx, sh, row):
    offs, pos = corpora.csr([np.array([1, 2, 3], np.uint32), np.array(row, np.uint32)])
    with pytest.raises(sh.StormHipError, match="sorted, distinct"):
        ctx.from_lists(offs, pos, 100, bitmap_policy=BUILD_ALL)
    # the context stays usable
    offs, pos = corpora.csr([np.array([1, 2, 3], np.uint32), np.array([2, 3], np.uint32)])
    with ctx.from_lists(offs, pos, 100, bitmap_policy=BUILD_ALL) as c:
        assert c.allpairs(policy="analytic")["sum"] == 2


def test_stormbin_ingest(ctx, sh, tmp_path):
    name, offs, pos, u = _generated(sh)[0]
    # empty rows in the file are skipped by the loader
    rows = [pos[offs[i]:offs[i + 1]] for i in range(len(offs) - 1)]
    rows = rows[:100] + [np.zeros(0, np.uint32)] * 5 + rows[100:]
    fo, fp = corpora.csr(rows)
    path = str(tmp_path / "z.bin")
    sh.write_stormbin(path, fo, fp, u)
    for max_rows, stride in ((0, 1), (0, 3), (250, 2)):
        print("stormbin", max_rows, stride, flush=True)
        eo, ep, eu = sh.stormbin_read(path, max_rows, stride)
        assert eu == u
        with ctx.from_stormbin(path, max_rows, stride, bitmap_policy=BUILD_ALL) as d, \
                ctx.from_lists(eo, ep, u, bitmap_policy=BUILD_ALL, on_host=True) as h:
            assert d.n == len(eo) - 1
            _same_arenas(d.arenas(), h.arenas(), (max_rows, stride))
            want = O.lists_closed_forms(eo, ep, u)[0]
            assert d.allpairs(policy="analytic", tile=64)["sum"] == want


def test_device_resident_lists(ctx, sh):
    name, offs, pos, u = _generated(sh)[1]
    # offsets need not start at 0: the corpus takes rows 5.. of the buffers
    with sh.DeviceArray(offs[5:].astype(np.uint64)) as do, \
            sh.DeviceArray(pos.astype(np.uint32)) as dp:
        with ctx.from_device_lists(do, dp, u, bitmap_policy=BUILD_ALL) as d, \
                ctx.from_lists(offs[5:] - offs[5], pos[offs[5]:], u, bitmap_policy=BUILD_ALL,
                               on_host=True) as h:
            _same_arenas(d.arenas(), h.arenas(), "device lists")
            st = d.allpairs(policy="analytic", tile=128)
            want = O.lists_closed_forms(offs[5:] - offs[5], pos[offs[5]:], u)[0]
            assert st["sum"] == want
