"""CPU checks of the accel.h surface (include/accel_surface.h): the library
exports every declared function, its structs have the reference's layout,
and the list functions (sort_accelcands, eliminate_harmonics) order and
prune exactly as the oracle's restatement of src/accel_utils.c does.  No
GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import accel as A, surface as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/include"


def _declared():
    txt = open(os.path.join(ROOT, "include", "accel_surface.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w \*]*?\b([a-z_][a-z_0-9]*)\s*\(", txt,
                                 flags=re.M)))


def test_surface_exports_every_declared_function():
    lib = C.CDLL(A.LIB_PATH)
    names = _declared()
    assert len(names) >= 29
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(S.EXPORTED)


def _ctypes_layout(st):
    return [C.sizeof(st)] + [getattr(st, f[0]).offset for f in st._fields_]


@pytest.mark.parametrize("name", sorted(S.LAYOUT_TYPES))
def test_struct_layout_matches_python_mirror(name):
    assert S.layout(name) == _ctypes_layout(S.LAYOUT_TYPES[name])


# --- the reference's own header text, parsed (this container only) ---------

_CT = {"long long": C.c_longlong, "int": C.c_int, "float": C.c_float, "double": C.c_double,
       "char": C.c_char, "unsigned short": C.c_ushort, "long": C.c_long}


def _parse_struct(text, name, macros):
    end = re.search(r"\}\s*" + name + r"\s*;", text)
    assert end, name
    start = text.rfind("typedef struct", 0, end.start())
    body = text[text.index("{", start) + 1:end.start()]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    body = re.sub(r"//[^\n]*", "", body)
    fields = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        parts = decl.split(",")
        dm = re.match(r"^(.*?)([\*\s]*\w+\s*(?:\[[^\]]*\])?)$", parts[0])
        base = dm.group(1).replace("struct ", "").strip()
        for d in [dm.group(2)] + parts[1:]:
            am = re.match(r"^([\*\s]*)(\w+)\s*(?:\[(.*)\])?$", d.strip())
            typ = C.c_void_p if "*" in am.group(1) else _CT[base]
            if am.group(3):
                n = eval(am.group(3), {}, dict(macros))
                fields.append((am.group(2), typ * n))
            else:
                fields.append((am.group(2), typ))
    return type(name, (C.Structure,), {"_fields_": fields})


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference headers not present")
@pytest.mark.parametrize("name,header", [
    ("accelobs", "accel.h"), ("accelcand", "accel.h"), ("kernel", "accel.h"),
    ("subharminfo", "accel.h"), ("ffdotpows", "accel.h"), ("infodata", "makeinf.h"),
    ("Cmdline", "accelsearch_cmd.h"), ("rderivs", "presto.h"), ("fourierprops", "presto.h")])
def test_struct_layout_matches_reference_header(name, header):
    text = open(os.path.join(REF, header)).read()
    macros = {k: int(v) for k, v in re.findall(r"#define\s+(\w+)\s+(\d+)\b", text)}
    ref = _parse_struct(text, name, macros)
    assert S.layout(name) == _ctypes_layout(ref)


# --- list functions against the oracle ---------------------------------------

_libc = C.CDLL(None)
_libc.calloc.restype = C.c_void_p
_libc.calloc.argtypes = [C.c_size_t, C.c_size_t]


def _glib_loaded():
    return hasattr(_libc, "g_slist_alloc")


def _make_list(items):
    """A GSList of accelcand in C heap memory, as the library would build."""
    head = C.POINTER(S.GSList)()
    for it in reversed(items):
        node = C.cast(_libc.calloc(1, C.sizeof(S.GSList)), C.POINTER(S.GSList))
        cand = C.cast(_libc.calloc(1, C.sizeof(S.AccelCand)), C.POINTER(S.AccelCand))
        cand.contents.power, cand.contents.sigma = it["power"], it["sigma"]
        cand.contents.numharm, cand.contents.r = it["numharm"], it["r"]
        node.contents.data = C.cast(cand, C.c_void_p)
        node.contents.next = head
        head = node
    return head


def _read(head):
    out, p = [], head
    while p:
        c = C.cast(p.contents.data, C.POINTER(S.AccelCand)).contents
        out.append((c.sigma, c.power, c.r))
        p = p.contents.next
    return out


def _free(head):
    p = head
    while p:
        S.lib.free_accelcand(p.contents.data, None)
        p = p.contents.next
    S.lib.accel_gslist_free(head)


def _random_items(rng, n):
    items = []
    base = rng.uniform(50, 400)
    for _ in range(n):
        kind = rng.integers(4)
        if kind == 0:
            r = base * rng.integers(1, 6) + rng.uniform(-1, 1)
        elif kind == 1:
            r = base * rng.choice([1.5, 2.5, 2 / 3, 0.75, 0.4]) + rng.uniform(-1, 1)
        else:
            r = rng.uniform(20, 5000)
        # repeated sigmas exercise the power tie-break and the stability
        sig = float(np.float32(rng.choice([3.0, 4.5, rng.uniform(2, 12)])))
        items.append(dict(power=float(np.float32(rng.uniform(20, 80))), sigma=sig,
                          numharm=int(rng.choice([1, 2, 4, 8])), r=float(r)))
    return items


def _make_olist(items):
    """The same list as the oracle's o_cand chain, in C heap memory."""
    head = C.POINTER(O.OCand)()
    for it in reversed(items):
        node = C.cast(_libc.calloc(1, C.sizeof(O.OCand)), C.POINTER(O.OCand))
        node.contents.power, node.contents.sigma = it["power"], it["sigma"]
        node.contents.numharm, node.contents.r = it["numharm"], it["r"]
        node.contents.next = head
        head = node
    return head


def _read_o(head):
    out, p = [], head
    while p:
        c = p.contents
        out.append((c.sigma, c.power, c.r))
        p = p.contents.next
    return out


@pytest.mark.skipif(_glib_loaded(), reason="glib in the process: nodes come from glib")
@pytest.mark.parametrize("seed", range(8))
def test_sort_and_eliminate_match_oracle(seed):
    rng = np.random.default_rng(seed)
    items = _random_items(rng, 60)
    ours = S.lib.sort_accelcands(_make_list(items))
    ref = O.lib.o_sort_accelcands(_make_olist(items))
    assert _read(ours) == _read_o(ref)
    n, m = C.c_int(len(items)), C.c_int(len(items))
    ours = S.lib.eliminate_harmonics(ours, C.byref(n))
    ref = O.lib.o_eliminate_harmonics(ref, C.byref(m))
    assert _read(ours) == _read_o(ref)
    assert n.value == m.value < len(items)
    _free(ours)
    O.lib.o_free_cands(ref)
