"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and rejects bad arguments with status codes before touching a GPU."""
import ctypes
import os
import re

import pytest

import oracle
import vjoin

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vj.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vj_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("vj_hash", "vj_join", "vj_join_into", "vj_ctx_create", "vj_partition", "vj_join_batch"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = vjoin.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings():
    L = vjoin.lib()
    for code in vjoin.STATUS:
        s = L.vj_status_str(code)
        assert s and s.decode() != "unknown status"
    assert L.vj_status_str(999) == b"unknown status"


def test_hash_names_match_oracle():
    L = vjoin.lib()
    for k, name in enumerate(vjoin.HASH_KINDS):
        assert vjoin.kind_id(name) == k == oracle.kind_id(name)
        assert L.vj_hash_kind_name(k).decode() == name
    assert L.vj_hash_kind_name(13) is None
    k = ctypes.c_int32()
    # "xor_hw" appears in the reference TOMLs but matches no kernel (hash.cu:515-541)
    assert L.vj_hash_kind_from_name(b"xor_hw", ctypes.byref(k)) == vjoin.STATUS_CODE["VJ_ERR_UNKNOWN_HASH"]


def _empty_table(n_cols=1, key_width=8, n_rows=0):
    t = vjoin.Table()
    t.n_rows = n_rows
    t.n_cols = n_cols
    t.key_width = key_width
    return t


@pytest.mark.parametrize("mutate,code", [
    (lambda t, o: setattr(t, "n_cols", 0), "VJ_ERR_INVALID_ARGUMENT"),
    (lambda t, o: setattr(t, "n_cols", 5), "VJ_ERR_INVALID_ARGUMENT"),
    (lambda t, o: setattr(t, "key_width", 2), "VJ_ERR_INVALID_ARGUMENT"),
    (lambda t, o: setattr(o, "kind", 13), "VJ_ERR_UNKNOWN_HASH"),
    (lambda t, o: setattr(o, "kind", -1), "VJ_ERR_UNKNOWN_HASH"),
    (lambda t, o: setattr(o, "hash_bits", 16), "VJ_ERR_INVALID_ARGUMENT"),
    (lambda t, o: setattr(o, "chunk_bits", 16), "VJ_ERR_INVALID_ARGUMENT"),
    (lambda t, o: setattr(o, "chunk_bits", 128), "VJ_ERR_INVALID_ARGUMENT"),  # 1 column = 8 bytes < 16-byte chunk
    (lambda t, o: setattr(t, "n_rows", 1 << 32), "VJ_ERR_UNSUPPORTED"),
    # one row past the limit is refused before anything else is looked at
    (lambda t, o: setattr(t, "n_rows", (1 << 32) - (1 << 16) + 1), "VJ_ERR_UNSUPPORTED"),
    # at the limit the row count is accepted; the NULL key column is what fails
    (lambda t, o: setattr(t, "n_rows", (1 << 32) - (1 << 16)), "VJ_ERR_INVALID_ARGUMENT"),
])
def test_hash_argument_errors(mutate, code):
    t = _empty_table()
    o = vjoin.HashOpts()
    vjoin.lib().vj_hash_opts_default(ctypes.byref(o))
    mutate(t, o)
    rc = vjoin.lib().vj_hash(ctypes.byref(t), ctypes.byref(o), None, None)
    assert vjoin.STATUS[rc] == code


def test_hash_empty_table_is_ok_without_device():
    t = _empty_table()
    o = vjoin.HashOpts()
    vjoin.lib().vj_hash_opts_default(ctypes.byref(o))
    assert (o.kind, o.hash_bits, o.chunk_bits, o.word_mode) == (0, 64, 64, 0)
    assert vjoin.lib().vj_hash(ctypes.byref(t), ctypes.byref(o), None, None) == 0
    assert vjoin.lib().vj_hash(None, ctypes.byref(o), None, None) == vjoin.STATUS_CODE["VJ_ERR_INVALID_ARGUMENT"]


def test_row_limit_matches_header():
    text = open(HEADER).read()
    m = re.search(r"#define VJ_MAX_ROWS (\d+)ull", text)
    assert m and int(m.group(1)) == (1 << 32) - (1 << 16)


# ctypes mirrors in vjoin.py and the C structs they stand for
MIRRORS = {"Table": "vj_table", "HashOpts": "vj_hash_opts", "JoinOpts": "vj_join_opts", "Rows": "vj_rows",
           "Pairs": "vj_pairs", "Request": "vj_request", "Response": "vj_response",
           "ServerStats": "vj_server_stats", "JoinSummary": "vj_join_summary"}


def test_ctypes_mirrors_match_the_header(tmp_path):
    """Every ctypes mirror has the size and field offsets of its C struct, as gcc lays it out."""
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "vj.h"', "int main(void) {"]
    want = []
    for py, c in MIRRORS.items():
        cls = getattr(vjoin, py)
        lines.append('printf("%%zu\\n", sizeof(%s));' % c)
        want.append(ctypes.sizeof(cls))
        for name, _ in cls._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (c, name))
            want.append(getattr(cls, name).offset)
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want
