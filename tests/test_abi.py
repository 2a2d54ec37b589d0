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
                   "int main(){ storm::gpu::AllPairsStats s; (void)s;\n"
                   "  auto f = &storm::gpu::allpairs_sum; auto g = &storm::gpu::allpairs_tiles;\n"
                   "  (void)f; (void)g; return storm::gpu::name_of(storm::gpu::Policy::PerTile)[0] == 'p' ? 0 : 1; }\n")
    rc = os.system(f"g++ -std=c++17 -Wall -Werror -I{ROOT}/include {src} -o {tmp_path}/t "
                   f"-L{PKG} -lstorm_gpu -lstorm_hip -Wl,-rpath,{PKG}")
    assert rc == 0
    assert os.system(f"{tmp_path}/t") == 0
