"""The compressed coefficient layout on the host (no GPU): the per-cell
compression the engine applies (ems_compress_operator) and the import of the
reference's SSE-compressed operator (ems_expand_sse_compressed), checked
against the operator arrays themselves and the oracle's restatement of
Operator_SSE_Compressed::CompressOperator."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from openems_amd import scenes
from openems_amd._native import NativeError
from openems_amd.engine import compress_operator, expand_sse_compressed
from openems_amd.operator import Operator
from oracle.oracle import RefEngine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _operator(scene):
    op = Operator(scene, nthreads=4)
    return op, [op.coeff(n) for n in ("vv", "vi", "ii", "iv")]


def _cells_xyz_first(a):
    """[3][Nx][Ny][Nz] -> rows of cells in x -> y -> z order (x outermost)"""
    return a.reshape(3, -1).T


@pytest.fixture(scope="module")
def bricks():
    op, arrs = _operator(scenes.upml_bricks(40, 100, nbricks=8, dump=False))
    yield op, arrs
    op.close()


def test_tables_reconstruct_every_coefficient(native, bricks):
    op, (vv, vi, ii, iv) = bricks
    N = op.N
    idx, tv, ti = compress_operator(N, vv, vi, ii, iv)
    # idx is [Nz][Ny][Nx]; bring it to [Nx][Ny][Nz]
    u = idx.transpose(2, 1, 0)
    for n in range(3):
        assert np.array_equal(tv[u, 2 * n], vv[n]) and np.array_equal(tv[u, 2 * n + 1], vi[n])
        assert np.array_equal(ti[u, 2 * n], ii[n]) and np.array_equal(ti[u, 2 * n + 1], iv[n])
    assert np.all(tv[:, 6:] == 0) and np.all(ti[:, 6:] == 0)


def test_first_seen_numbering_and_unique_count(native, bricks):
    op, (vv, vi, ii, iv) = bricks
    idx, tv, ti = compress_operator(op.N, vv, vi, ii, iv)
    order = idx.transpose(2, 1, 0).reshape(-1)          # cells x -> y -> z
    first = np.maximum.accumulate(order)
    # every index is either seen before or the next new number
    new = np.concatenate([[True], order[1:] > first[:-1]])
    assert order[0] == 0
    assert np.array_equal(order[new], np.arange(new.sum()))
    rows = np.stack([_cells_xyz_first(a) for a in (vv, vi, ii, iv)], axis=1).reshape(-1, 12) + np.float32(0)
    U = np.unique(rows.view(np.dtype((np.void, 48))).ravel()).size
    assert len(tv) == U == new.sum()
    assert U < op.num_cells / 10


def test_negative_zero_is_the_same_key(native):
    N = (4, 4, 4)
    vv = np.ones((3,) + N, np.float32)
    vi = np.zeros((3,) + N, np.float32)
    vi[1, 2, 3, 1] = -0.0
    idx, tv, _ = compress_operator(N, vv, vi, vv, vi)
    assert len(tv) == 1 and np.all(idx == 0)
    assert not np.signbit(tv[0, 3])                    # the first-seen entry is stored


def test_too_many_unique_tuples_is_refused(native):
    N = (48, 48, 36)
    rng = np.random.default_rng(5)
    arrs = [rng.random((3,) + N, dtype=np.float32) for _ in range(4)]
    with pytest.raises(NativeError):
        compress_operator(N, *arrs)


def test_sse_compressed_import_reproduces_operator(native, bricks):
    """The oracle's SSE-compressed operator (4 z-lanes per entry, lane l holds
    z = l*nV + v) expands back to the operator it was compressed from."""
    op, arrs = bricks
    from oracle.operator_ref import RefOperator
    ref = RefOperator(scenes.upml_bricks(40, 100, nbricks=8, dump=False)).build()
    eng = RefEngine(ref, "sse_compressed")
    opi = eng.op_index_full()
    tabs = eng.ctables()
    assert eng.num_unique() == len(tabs[0][0])
    out = expand_sse_compressed(op.N, opi, *tabs)
    eng.close()
    for got, want in zip(out, arrs):
        assert np.array_equal(got, want)


def test_sse_compressed_import_rejects_bad_index(native):
    N = (4, 4, 5)
    opi = np.zeros(N, np.uint32)
    opi[1, 1, 1] = 3
    tab = [np.ones((2, 4), np.float32)] * 3
    with pytest.raises(NativeError):
        expand_sse_compressed(N, opi, tab, tab, tab, tab)


def test_update_kernels_have_no_fused_or_packed_fp32(native):
    """Bit-exactness needs every product and sum rounded on its own: the
    gfx950 code of the update kernels (both half steps and both single-sweep
    steps) holds no f32 FMA / MAC and no packed f32 instruction."""
    src = os.path.join(ROOT, "openems_amd", "csrc", "engine", "ems_engine.hip")
    from openems_amd import build as B
    with tempfile.TemporaryDirectory() as d:
        asm = os.path.join(d, "engine.s")
        cmd = [B.HIPCC] + B.device_flags() + ["--cuda-device-only", "-S", "-o", asm, src]
        subprocess.run(cmd, check=True, capture_output=True)
        text = open(asm).read()
    kernels = {}
    cur = None
    for line in text.splitlines():
        m = re.match(r"^(_Z\S+):", line)
        if m:
            cur = m.group(1)
            kernels.setdefault(cur, [])
        elif cur:
            kernels[cur].append(line)
    upd = {k: v for k, v in kernels.items() if re.search(r"k_(volt|curr|step|step2)I", k)}
    # 3 UPML modes x 2 layouts x (2 half steps x 2 field-access forms + 2 single-sweep forms)
    assert len(upd) == 36
    bad = re.compile(r"\bv_(fma\w*_f32|fmac\w*_f32|mac_f32|mad_f32|pk_\w+_f32|fma_mix\w*|dot\w*)")
    for k, lines in upd.items():
        hits = [l.strip() for l in lines if bad.search(l)]
        assert not hits, f"{k}: {hits[:4]}"
