"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import os
import re

from hspz_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(only=None):
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h") or (only and h != only):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(hspz_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    lib = E.lib()
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_export_list_matches_header():
    from hspz_amd import seqio
    assert set(E.EXPORTS) | set(seqio.EXPORTS) <= declared_symbols()
    assert set(seqio.EXPORTS) == declared_symbols("hspz_io.h")


def test_device_count_is_callable_without_gpu():
    assert E.device_count() >= 0


def test_flag_and_stat_tables_match_header():
    text = open(os.path.join(ROOT, "include", "hspz.h")).read()
    found = re.findall(r"#define HSPZ_(FLAG|STAT)_([A-Z0-9_]+)\s+(\d+)", text)
    flags = {name.lower(): int(v) for kind, name, v in found if kind == "FLAG"}
    stats = {name.lower(): int(v) for kind, name, v in found if kind == "STAT"}
    assert flags == E.FLAG
    assert stats == E.STAT


def test_list_capacity_growth_is_bounded():
    """The batched pass's overflow rerun grows its survivor and deferred-hit
    lists in 64-bit arithmetic: to the larger of the count reached + 1/4 and
    1.5 x the old capacity, clamped to 2^30 records; a count beyond 2^30 can
    never be held and is refused (the pass then fails with HSPZ_E_OOM instead
    of rerunning forever)."""
    import ctypes as C
    f = E.lib().hspz_test_grow_list_cap
    f.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    out = C.c_uint32()
    cases = [  # cap, count, holds, new cap
        (64, 100, 1, 125),
        (1 << 22, 5, 1, 3 << 21),
        (1 << 28, 300_000_000, 1, 3 << 27),
        (1 << 20, 300_000_000, 1, 375_000_000),
        (1 << 29, (1 << 29) + 1, 1, 3 << 28),
        ((1 << 30) - 1, 1 << 30, 1, 1 << 30),  # 1.25 x 2^30 would pass the clamp
        (1 << 31, (1 << 31) + 1, 0, None),
        (1 << 30, (1 << 30) + 1, 0, None),
        (0xFFFFFFFF, 0xFFFFFFFF, 0, None),
    ]
    for cap, count, holds, want in cases:
        got = f(cap, count, C.byref(out))
        assert got == holds, (cap, count)
        if holds:
            assert out.value == want and out.value >= count, (cap, count, out.value)
