"""The C-ABI library loads and exports every entry point include/*.h
declares (no GPU needed: nothing is called)."""
import ctypes
import glob
import os
import re

from gi_test_util import ROOT, gi


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(gi_\w+)\s*\(", text, flags=re.M))
    return names


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 15
    lib = ctypes.CDLL(gi.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(gi.SYMBOLS) <= names


def test_status_strings_and_null_arguments():
    lib = gi.lib()
    assert lib.gi_status_string(0) == b"ok"
    assert lib.gi_destroy(None) == 1  # GI_ERR_INVALID_ARGUMENT
    assert lib.gi_intersect(None, None, None) == 1
    assert lib.gi_result_free(None) == 1
    # serving entry points reject missing arguments before touching a device
    t = ctypes.c_uint64()
    assert lib.gi_server_create(None, None, 0, None, 0, None) == 1
    assert lib.gi_server_submit(None, 0, None, ctypes.byref(t)) == 1
    assert lib.gi_server_wait(None, 1, None, None) == 1
    assert lib.gi_server_destroy(None) == 1
    assert lib.gi_share_mesh(None, 0, None, 0) == 1


def test_host_exact_stage_matches_golden():
    import json
    from fractions import Fraction
    cases = json.load(open(os.path.join(ROOT, "tests", "golden", "tritri_cases.json")))["cases"]
    for c in cases:
        a = [float.fromhex(x) for x in c["a"]]
        b = [float.fromhex(x) for x in c["b"]]
        hit, cop, segs = gi.exact_pair_host(a, b)
        want = sorted(tuple(tuple(Fraction(*map(int, x.split("/"))) for x in P) for P in s) for s in c["segments"])
        assert hit == c["intersect"], c["name"]
        assert segs == want, c["name"]
