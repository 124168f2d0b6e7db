"""The C ABI library loads and exports every entry point include/tess.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

import tess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tess_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared("tess.h") == sorted(tess.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tess.LIB_PATH)
    missing = [s for s in declared("tess.h") if not hasattr(lib, s)]
    assert not missing, missing


def test_opts_default_without_device():
    lib = tess.load_library()
    o = tess.Opts()
    lib.tess_opts_default(ctypes.byref(o))
    assert o.cell_density == 1.25 and o.mode == tess.MODE_PLAIN and o.vertex_capacity == 0


def test_null_arguments_are_rejected():
    lib = tess.load_library()
    assert lib.tess_diagram_compute(None) == 1  # TESS_E_INVALID
    assert lib.tess_ctx_create(0, None) == 1


def _xoshiro_first(seed, k):
    M = (1 << 64) - 1

    def splitmix(x):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return x, z ^ (z >> 31)

    s, x = [], seed
    for _ in range(4):
        x, v = splitmix(x)
        s.append(v)
    rotl = lambda v, r: ((v << r) | (v >> (64 - r))) & M  # noqa: E731
    out = []
    for _ in range(k):
        r = (rotl((s[1] * 5) & M, 7) * 9) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        out.append((r >> 11) * 2.0 ** -53)
    return out


def test_uniform_generator_is_xoshiro256ss():
    pts = tess.uniform(5, 1234)
    assert list(pts.reshape(-1)) == _xoshiro_first(1234, 15)


def test_header_is_c99_and_struct_sizes_match(tmp_path):
    """include/tess.h compiles as plain C99, and the binding's structures have
    the sizes the C compiler gives them."""
    import subprocess

    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "tess.h"\nint main(void) {\n'
                   '  printf("%zu %zu %zu %zu\\n", sizeof(tess_opts), sizeof(tess_cell_stats), '
                   'sizeof(tess_timings), sizeof(tess_batch_config));\n  return 0;\n}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I",
                    os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == [ctypes.sizeof(tess.Opts), ctypes.sizeof(tess.CellStats), ctypes.sizeof(tess.Timings),
                   ctypes.sizeof(tess.BatchConfig)]
