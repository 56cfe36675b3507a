"""The C-ABI library: loads, exports every symbol include/albatros_hip.h
declares, and the ctypes table matches the header (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "albatros_hip.h")


def header_symbols():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|const char\*)\s+(alb_\w+)\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert len(syms) >= 25
    assert "alb_repfb" in syms and "alb_xcorr_all_ant" in syms and "alb_unpack_4bit_float" in syms


def test_library_exports_every_header_symbol():
    from albatros_axion_detection_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from albatros_axion_detection_amd import _lib
    assert sorted(_lib.SIGNATURES) == header_symbols()
    txt = open(HDR).read().replace("\n", " ")
    for name, (_, args) in _lib.SIGNATURES.items():
        m = re.search(r"\b%s\(([^)]*)\)" % name, txt)
        decl = m.group(1).strip()
        n = 0 if decl in ("", "void") else decl.count(",") + 1
        assert n == len(args), (name, n, len(args))


def test_library_loads_and_reports_version():
    from albatros_axion_detection_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.alb_version() >= 100
    assert lib.alb_repfb_workspace_bytes(2, 64) == 2 * 32 * 4096 * 8


def test_invalid_arguments_fail_loudly_without_gpu_work():
    # argument checks run before any launch: null pointers are rejected
    from albatros_axion_detection_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    with pytest.raises(ValueError):
        _lib.call("alb_sortpols", None, None, None, 0, 1, 1, 8, 4, 0, 1, None)
    with pytest.raises(ValueError):
        _lib.call("alb_xcorr_all_ant", None, 8, 16, 4, None, 1.0, None)


def test_repfb_interior_blocks_brute_force():
    # the interior-block range of a tile against a direct scan (host-only entry point)
    from albatros_axion_detection_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    import numpy as np
    rng = np.random.default_rng(7)
    lo, hi = ctypes.c_int64(), ctypes.c_int64()
    for _ in range(400):
        R1 = int(rng.choice([16, 32, 64]))
        ntap = int(rng.integers(1, 6))
        nkeep = int(rng.integers(0, 600))
        g0 = int(rng.integers(0, 1000))
        nblocks = int(rng.integers(0, 40))
        _lib.call("alb_repfb_interior", g0, nkeep, R1, ntap, nblocks, ctypes.byref(lo), ctypes.byref(hi))
        want = [b for b in range(nblocks) if R1 * b >= g0 and R1 * (b + ntap) <= g0 + nkeep]
        got = list(range(lo.value, hi.value))
        assert got == want, (g0, nkeep, R1, ntap, nblocks)


def test_fused_column_pass_shape_probe():
    # ntiles = 0 checks the shape only (no pointers, no GPU work)
    from albatros_axion_detection_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    probe = lambda nslice, cut, ntap, osamp: lib.alb_ipfb_columns_repfb(
        None, 4096, 0, 0, nslice, None, nslice // 2 + 1, cut, None, 0, nslice - 2 * cut, None, ntap, osamp, 0, None,
        None)
    assert probe(4096, 4, 4, 64) == 0
    assert probe(4096, 4, 4, 32) == 0
    assert probe(4096, 4, 4, 16) == 3      # ENOTSUP: falls back to the two-kernel form
    assert probe(2048, 4, 4, 64) == 3
    assert probe(4096, 4, 9, 64) == 3      # window strip too large
    with pytest.raises(ValueError):
        _lib.call("alb_ipfb_columns_repfb", None, 4096, 0, 1, 4096, None, 2049, 4, None, 0, 4088, None, 4, 64, 10,
                  None, None)


def test_split_column_pass_shape_probe_and_refusals():
    # the split form (tile-boundary blocks in two halves): shapes it takes,
    # and undersized boundary arrays / bad geometry refused before any launch
    from albatros_axion_detection_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    probe = lambda nslice, cut, ntap, osamp: lib.alb_ipfb_columns_repfb_split(
        None, 4096, 0, 0, nslice, None, nslice // 2 + 1, cut, 0, nslice - 2 * cut, None, ntap, osamp, 0, None, None,
        0, 0, None)
    assert probe(4096, 4, 4, 64) == 0
    assert probe(4096, 10, 3, 64) == 0
    assert probe(4096, 4, 4, 32) == 3      # ENOTSUP: osamp 64 only
    assert probe(4096, 0, 4, 64) == 3      # needs cut >= 1 (the zeroed cut rows)
    assert probe(4096, 4, 5, 64) == 3      # ntap <= 4
    assert probe(2048, 4, 4, 64) == 3
    assert lib.alb_repfb_split_slot_bytes(64) == lib.alb_repfb_workspace_bytes(1, 64)
    fake = ctypes.c_void_p(1 << 20)        # never dereferenced: the checks fail first
    # tiles 3..4 of a window need ntap * 4 = 16 slots, 15 given
    with pytest.raises(ValueError):
        _lib.call("alb_ipfb_columns_repfb_split", fake, 8, 4096 * 4096, 2, 4096, fake, 2049, 4, 3 * 4088, 4088, fake,
                  4, 64, 1000, fake, fake, 3, 15, None)
    # the window's tile 0 must start at or before row 0: tile0 0 at row 100
    # would put the blocks before row 100 in a slot before P
    with pytest.raises(ValueError, match="after row 0"):
        _lib.call("alb_ipfb_columns_repfb_split", fake, 8, 4096 * 4096, 2, 4096, fake, 2049, 4, 100, 4088, fake,
                  4, 64, 1000, fake, fake, 0, 16, None)
    with pytest.raises(ValueError, match="after row 0"):
        _lib.call("alb_ipfb_columns_repfb_spectral_split", fake, 2, 4096, fake, 2049, 4, 2 * 4088 + 1, 4088, fake,
                  4, 64, 1000, fake, fake, 2, 16, None)
    # consecutive tiles must be nslice - 2 cut rows apart
    with pytest.raises(ValueError):
        _lib.call("alb_ipfb_columns_repfb_split", fake, 8, 4096 * 4096, 2, 4096, fake, 2049, 4, 0, 4000, fake,
                  4, 64, 1000, fake, fake, 0, 16, None)
    # stage 2: osamp 64 only, g_first <= 0, blocks inside the tiles
    stage2 = lambda osamp, nblock, g_first, ntiles: _lib.call(
        "alb_repfb_stage2_split", fake, 0, fake, 0, 1, nblock, osamp, 64000, 83200, fake, 1, 4, 4, 1024, 4096,
        g_first, 4088, ntiles, 4, None)
    with pytest.raises(ValueError):
        stage2(32, 10, 0, 2)
    with pytest.raises(ValueError):
        stage2(64, 10, 5, 2)
    with pytest.raises(ValueError):
        stage2(64, 130, 0, 2)                # block 129 starts past two tiles of 4088 rows


def test_split_halves_cover_every_block_once():
    # The split column pass's block ranges and boundary slots, and stage 2's
    # slot lookup (the library's own index functions, host-side), over random
    # stream windows as the batched pass plans them: every block with data
    # gets exactly its rows below the end of the data, as A (from the tile
    # holding its first row) plus at most one boundary half (from the next
    # tile); no destination is written twice and every written half is read.
    from albatros_axion_detection_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    import numpy as np
    rng = np.random.default_rng(11)
    lo, hi, e0 = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    R1, nwin = 64, 0
    for _ in range(120):
        ntap = int(rng.integers(1, 5))
        cut = int(rng.integers(1, 40))
        acclen = 4096 - 2 * cut
        nblock = int(rng.choice([3, 5, 16, 64, 1024]))
        nchunks = int(rng.integers(2, 30))
        njobs = max(1, (nchunks * acclen // R1 - (ntap - 1)) // nblock)
        s0 = int(rng.integers(0, njobs))
        s1 = int(rng.integers(s0 + 1, njobs + 1))
        nvalid = int(rng.integers(1, nchunks + 1))
        row_lo, row_hi = s0 * nblock * R1, s1 * nblock * R1 + (ntap - 1) * R1
        c_lo = min(nvalid, row_lo // acclen)
        c_hi = max(c_lo, min(nvalid, -(-row_hi // acclen)))
        ntiles, nbg = c_hi - c_lo, (s1 - s0) * nblock
        if ntiles == 0:
            continue
        g_first = c_lo * acclen - row_lo
        rows_valid = c_hi * acclen - row_lo
        nbz = min(nbg, -(-rows_valid // R1))
        written, owner = {}, {}
        for tt in range(ntiles):
            g0 = g_first + tt * acclen
            _lib.call("alb_repfb_split_blocks", g0, acclen, R1, ntap, nbg, ctypes.byref(lo), ctypes.byref(hi),
                      ctypes.byref(e0))
            for b in range(lo.value, hi.value):
                rows = {r for r in range(R1 * b, R1 * (b + ntap)) if g0 <= r < g0 + acclen}
                assert rows, (b, g0)
                dst = ("A", b) if R1 * b >= g0 else ("P", ntap * (tt - 1) + b - e0.value)
                if dst[0] == "P":
                    assert tt >= 1 and 0 <= dst[1] < ntap * ntiles
                assert dst not in written
                written[dst] = rows
                owner[dst] = b
        read = set()
        for b in range(nbz):
            rows = set(written.get(("A", b), set()))
            e = int(_lib.load().alb_repfb_split_slot(b, g_first, acclen, ntiles, ntap, R1))
            assert e >= -1
            if e >= 0:
                assert ("P", e) in written and not rows & written[("P", e)]
                rows |= written[("P", e)]
                read.add(e)
            assert rows == {r for r in range(R1 * b, R1 * (b + ntap)) if r < rows_valid}, (b, ntap, cut)
        # every half written for a block with data is added by stage 2
        assert {d[1] for d in written if d[0] == "P" and owner[d] < nbz} == read
        nwin += 1
    assert nwin > 50
