"""The library's host-side integer setup (delfem_amd/csrc/dfm_host.cpp: FE
pattern, multicolour and block-multicolour orderings, ILU(k) symbolic, level
sets, sliced-ELL plans, phases found in natural-order factors, phase-tiled
sweep plans) built with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer (tests/cpp/host_check.cpp),
run on small meshes: no sanitizer report, the tile plans' layout invariants
hold (every row once, in-phase dependencies at lower local levels inside the
tile, slots in the factor's column order, other-phase columns finished by
earlier launches), and the pattern, ordering and ILU(k) pattern equal the
oracle's bit for bit."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from delfem_amd import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "cpp", "host_check.cpp"), os.path.join(ROOT, "delfem_amd", "csrc", "dfm_host.cpp")]

ORDER = {"natural": 0, "multicolour": 2, "block_multicolour": 3}


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are missing")
    exe = str(tmp_path_factory.mktemp("hc") / "host_check")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "delfem_amd", "csrc"), *SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def run(exe, tmp_path, prob, lev, ordering, block_nodes=16, pack_rows=64):
    conn = np.ascontiguousarray(prob.conn, dtype=np.uint32)
    nelem, npe = conn.shape
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([prob.nnode, nelem, npe, lev, ORDER[ordering], block_nodes, pack_rows, prob.b],
                         dtype=np.int32).tobytes())
        f.write(conn.tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    raw = np.fromfile(out, dtype=np.int32)
    status, nnz, nf, ncol, nlL, nlU, ntL, ntU, tiL, tiU, nph = (int(v) for v in raw[:11])
    n = prob.nnode
    parts, o = {}, 11
    for name, size in (("rowptr", n + 1), ("col", nnz), ("order", n), ("frow", n + 1), ("fcol", nf), ("flev", nf),
                       ("phase_ptr", ncol + 1 if ncol else (nph + 1 if nph else 0))):
        parts[name] = raw[o:o + size]
        o += size
    assert o == raw.size
    parts.update(status=status, ncolour=ncol, nlev=(nlL, nlU), nt=(ntL, ntU), ntile=(tiL, tiU), nphase=nph)
    return parts


CASES = [
    ("poisson2d_20", lambda: meshgen.poisson2d(20, seed=1), 0, "natural", 16, 64),
    ("poisson2d_20", lambda: meshgen.poisson2d(20, seed=1), 2, "natural", 16, 64),
    ("poisson2d_20", lambda: meshgen.poisson2d(20, seed=1), 0, "multicolour", 16, 64),
    ("poisson2d_20", lambda: meshgen.poisson2d(20, seed=1), 1, "multicolour", 16, 64),
    ("poisson2d_20", lambda: meshgen.poisson2d(20, seed=1), 2, "block_multicolour", 16, 1),
    ("oseen2d_16", lambda: meshgen.oseen2d(16, seed=3), 1, "block_multicolour", 8, 64),
    ("oseen2d_16", lambda: meshgen.oseen2d(16, seed=3), 1, "block_multicolour", 48, 256),
    ("oseen2d_16", lambda: meshgen.oseen2d(16, seed=3), 2, "block_multicolour", 100, 64),
    ("solid3d_4", lambda: meshgen.solid3d(4, seed=3), 1, "multicolour", 16, 64),
    ("solid3d_4", lambda: meshgen.solid3d(4, seed=3), 1, "block_multicolour", 32, 128),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-ilu{c[2]}-{c[3]}-b{c[4]}-p{c[5]}" for c in CASES])
def test_host_setup_sanitized(host_check, tmp_path, orc, case):
    name, make, lev, ordering, bn, pack = case
    prob = make()
    got = run(host_check, tmp_path, prob, lev, ordering, bn, pack)
    assert got["status"] == 0
    rowptr, col = orc.fem_pattern(prob.nnode, prob.conn)
    assert np.array_equal(got["rowptr"], rowptr) and np.array_equal(got["col"], col)
    order = None
    if ordering == "multicolour":
        colour, nc, order = orc.colour_order(rowptr, col)
        assert got["ncolour"] == nc
        assert np.array_equal(got["phase_ptr"], np.concatenate([[0], np.cumsum(np.bincount(colour, minlength=nc))]))
    elif ordering == "block_multicolour":
        _, nc, order = orc.bmc_order(rowptr, col, bn)
        assert got["ncolour"] == nc
    if order is not None:
        assert np.array_equal(got["order"], order)
        pp = got["phase_ptr"]
        assert pp[0] == 0 and pp[-1] == prob.nnode and np.all(np.diff(pp) > 0)
    F = orc.iluk_symbolic(rowptr, col, lev, order)
    assert np.array_equal(got["frow"], F["rowptr"]) and np.array_equal(got["fcol"], F["col"])
    assert np.array_equal(got["flev"], F["lev"])
    if order is not None:
        # a colour ordering always has a tiling here (components stay far below 256 rows)
        assert got["nt"][0] in (64, 128, 256) and got["ntile"][0] > 0 and got["ntile"][1] > 0
        if pack == 1:
            # one component per tile: tiles = components >= rows / largest component
            assert got["ntile"][0] >= prob.nnode // got["nt"][0]


# natural-order factors: a large lexicographic numbering keeps its level sets
# (its phases would be long chains); a small one runs as a few 256-row tiles,
# one workgroup each, instead of one launch per level; a mesh numbered by
# colours of geometric tiles gives a few phases, far below the level count.
# Wherever phases are found, both sweeps have valid tile plans.
@pytest.mark.parametrize("name,make,lev,tiles", [
    ("poisson2d_160_lex", lambda: meshgen.poisson2d(160, seed=1), 0, False),
    ("poisson2d_20_lex", lambda: meshgen.poisson2d(20, seed=1), 2, True),
    ("oseen2d_32_lex", lambda: meshgen.oseen2d(32, seed=3), 1, True),
    ("oseen2d_32_tiles", lambda: meshgen.oseen2d(32, seed=3), 1, True),
    ("poisson2d_40_tiles", lambda: meshgen.poisson2d(40, seed=2), 2, True),
])
def test_phases_found(host_check, tmp_path, orc, name, make, lev, tiles):
    prob = make()
    if name.endswith("tiles"):
        prob = meshgen.relabel(prob, meshgen.tile_colour_order(prob, 16))
    got = run(host_check, tmp_path, prob, lev, "natural", 16, 64)
    print(name, got["nlev"], got["nphase"], got["phase_ptr"], got["ntile"], flush=True)
    assert got["status"] == 0
    rowptr, col = orc.fem_pattern(prob.nnode, prob.conn)
    F = orc.iluk_symbolic(rowptr, col, lev, None)
    assert np.array_equal(got["frow"], F["rowptr"]) and np.array_equal(got["fcol"], F["col"])
    if not tiles:
        assert got["nphase"] == 0 and got["ntile"] == (0, 0)
        return
    pp = got["phase_ptr"]
    assert 1 <= got["nphase"] <= max(got["nlev"]) // 4
    assert pp[0] == 0 and pp[-1] == prob.nnode and np.all(np.diff(pp) > 0)
    assert got["ntile"][0] > 0 and got["ntile"][1] > 0
