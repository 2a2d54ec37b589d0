"""The C ABI library loads and exports every symbol include/*.h declares; host
side error contract without a GPU (no compute calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stormbitmaps-internal_amd")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:storm_hip|STORM)_[A-Za-z0-9_]*)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("storm_hip.h", "libstorm_hip.so"),
                                        ("storm_gpu.h", "libstorm_gpu.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(PKG, lib)
    assert os.path.exists(path), f"{path} not built"
    L = C.CDLL(path, mode=C.RTLD_LOCAL)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_is_plain_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "storm_hip.h"\n#include "storm_gpu.h"\nint main(void){return 0;}\n')
    rc = os.system(f"gcc -std=c99 -Wall -Werror -I{ROOT}/include -c {src} -o {tmp_path}/t.o")
    assert rc == 0


def test_host_error_contract_without_device(sh):
    L = sh.lib()
    assert L.storm_hip_version().decode().startswith("storm_hip")
    assert L.storm_hip_strerror(-4) == b"out of memory"
    assert L.storm_hip_ctx_create(0, 0, None) == -1
    assert L.storm_hip_allpairs(None, None, None, None) == -1
    assert L.storm_hip_sync(None) == -1
    n = C.c_int(-1)
    assert L.storm_hip_device_count(C.byref(n)) == 0 and n.value >= 0
    if n.value == 0:
        h = C.c_void_p()
        assert L.storm_hip_ctx_create(0, 0, C.byref(h)) == -3     # device error, loudly
        assert b"device" in L.storm_hip_last_error()


def test_contig_dropin_null_contract():
    """storm.cpp:1176-1232 NULL-handle behaviour, no device touched."""
    L = C.CDLL(os.path.join(PKG, "libstorm_gpu.so"), mode=C.RTLD_LOCAL)
    L.STORM_contig_add.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    L.STORM_contig_clear.argtypes = [C.c_void_p]
    L.STORM_contig_pairw_intersect_cardinality.argtypes = [C.c_void_p]
    L.STORM_contig_pairw_intersect_cardinality.restype = C.c_uint64
    L.STORM_contig_new.restype = C.c_void_p
    L.STORM_contig_free.argtypes = [C.c_void_p]
    assert L.STORM_contig_add(None, None, 0) == -1
    assert L.STORM_contig_clear(None) == -1
    assert L.STORM_contig_pairw_intersect_cardinality(None) == 0
    h = L.STORM_contig_new(100)
    assert h
    assert L.STORM_contig_add(h, None, 1) == -2
    v = (C.c_uint32 * 2)(3, 200)
    assert L.STORM_contig_add(h, v, 2) == -3          # out of the vector length
    assert L.STORM_contig_add(h, v, 0) == 0           # nothing added
    assert L.STORM_contig_add(h, v, 1) == 1
    assert L.STORM_contig_pairw_intersect_cardinality(h) == 0   # one vector: no pairs
    L.STORM_contig_free(h)


def test_cpp_mirror_header_compiles_and_links(tmp_path):
    """include/storm_allpairs_gpu.h compiles as C++17 and its entry points are
    exported by libstorm_gpu.so (no device touched)."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "storm_allpairs_gpu.h"\n'
                   "using namespace storm::gpu;\n"
                   "int main(){ AllPairsStats s; (void)s;\n"
                   "  AllPairsStats (*f)(Engine&, const RowsCSR&, const storm_hip_cost_model*,"
                   " Policy, uint32_t, bool, Pairing) = &allpairs_sum;\n"
                   "  AllPairsStats (*g)(Engine&, storm_hip_corpus*, const storm_hip_cost_model*,"
                   " Policy, TileCallback, void*, uint32_t) = &allpairs_tiles;\n"
                   "  int (*h)(RowArenas&, uint32_t*) = &infer_universe;\n"
                   "  Engine& (*e)() = &default_engine;\n"
                   "  (void)f; (void)g; (void)h; (void)e;\n"
                   "  return name_of(Policy::PerTile)[0] == 'p' ? 0 : 1; }\n")
    rc = os.system(f"g++ -std=c++17 -Wall -Werror -I{ROOT}/include {src} -o {tmp_path}/t "
                   f"-L{PKG} -lstorm_gpu -lstorm_hip -Wl,-rpath,{PKG}")
    assert rc == 0
    assert os.system(f"{tmp_path}/t") == 0


def test_infer_universe_from_row_arenas(tmp_path):
    """The reference-typed overloads recover the universe a Row vector was
    built over from build_row's gates (storm_repr.cpp:139,188): zone map iff
    card*512 > U, complement iff 2*card > U with U - card positions."""
    src = tmp_path / "u.cpp"
    src.write_text(r'''
#include <array>
#include <cstdio>
#include "storm_allpairs_gpu.h"
using namespace storm::gpu;
// rows: (card, last, has_bitmap, has_occ, comp_len)
static int run(std::initializer_list<std::array<uint32_t, 5>> rows, uint32_t nw, uint32_t* U,
               bool* comp) {
    RowArenas a;
    a.n_rows = rows.size(); a.n_words = nw;
    for (auto* o : {&a.list_off, &a.run_off, &a.ewah_off, &a.bitmap_off, &a.rank_off, &a.occ_off,
                    &a.comp_off}) o->assign(1, 0);
    for (auto r : rows) {
        const uint32_t m[8] = {r[0], 1, nw, 1, r[0] ? 0u : 0xFFFFFFFFu, r[1], 0, r[3]};
        a.meta.insert(a.meta.end(), m, m + 8);
        a.list_off.push_back(0); a.run_off.push_back(0);
        a.ewah_off.push_back(a.ewah_off.back() + (r[2] ? 1 : 0));
        a.bitmap_off.push_back(a.bitmap_off.back() + (r[2] ? nw : 0));
        a.rank_off.push_back(0);
        a.occ_off.push_back(a.occ_off.back() + (r[3] ? 1 : 0));
        a.comp_off.push_back(a.comp_off.back() + r[4]);
    }
    const int rc = infer_universe(a, U);
    *comp = a.has_comp;
    return rc;
}
int main() {
    uint32_t U = 0; bool comp = false; int bad = 0;
    // A zone-mapped row with its last bit at 100 over 2 words: U = 101.
    bad += run({{3, 100, 1, 1, 0}}, 2, &U, &comp) != 0 || U != 101;
    // A complement of 28 beside a 100-element row fixes U = 128.
    bad += run({{3, 100, 1, 1, 0}, {100, 127, 1, 1, 28}}, 2, &U, &comp) != 0 || U != 128 || !comp;
    // Complements disagreeing on U: no universe fits.
    bad += run({{100, 127, 1, 1, 28}, {90, 127, 1, 1, 30}}, 2, &U, &comp) == 0;
    // No zone map on a bitmap row: U >= card * 512 (4 * 512 over 32 words).
    bad += run({{4, 2000, 1, 0, 0}}, 32, &U, &comp) != 0 || U != 2048;
    // ... which cannot hold over 2 words.
    bad += run({{4, 100, 1, 0, 0}}, 2, &U, &comp) == 0;
    // A sparse-only row past half density drops every complement.
    bad += run({{3, 100, 1, 1, 0}, {60, 110, 0, 0, 0}}, 2, &U, &comp) != 0 || U != 111 || comp;
    std::printf("%d bad\n", bad);
    return bad;
}
''')
    rc = os.system(f"g++ -std=c++17 -Wall -Werror -I{ROOT}/include {src} -o {tmp_path}/u "
                   f"-L{PKG} -lstorm_gpu -lstorm_hip -Wl,-rpath,{PKG}")
    assert rc == 0
    assert os.system(f"{tmp_path}/u") == 0
